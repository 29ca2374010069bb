#define PF_REAL float
#define PF_VAR 1
#include "pf_row.inc"
