/* Test stand-in for Postgres's miscadmin.h: work_mem and
 * CHECK_FOR_INTERRUPTS (which here counts its calls). */
#ifndef PG_STANDIN_MISCADMIN_H
#define PG_STANDIN_MISCADMIN_H

extern int work_mem;

void pg_standin_check_for_interrupts(void);
#define CHECK_FOR_INTERRUPTS() pg_standin_check_for_interrupts()

#endif
