#define PF_REAL double
#define PF_VAR 0
#include "pf_row.inc"
