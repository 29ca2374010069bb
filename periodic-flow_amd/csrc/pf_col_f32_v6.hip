#define PF_REAL float
#define PF_VAR 6
#include "pf_col.inc"
