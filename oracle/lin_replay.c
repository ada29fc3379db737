/* lin_replay.c -- ORACLE / TEST INFRASTRUCTURE ONLY (linear-chain replay; filled in with the linear engine). */
#include <stdint.h>
int orc_lin_placeholder(void) { return 0; }
