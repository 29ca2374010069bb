#define PF_REAL float
#define PF_VAR 2
#include "pf_col.inc"
