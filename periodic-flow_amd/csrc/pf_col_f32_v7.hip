#define PF_REAL float
#define PF_VAR 7
#include "pf_col.inc"
