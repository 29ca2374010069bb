#define PF_REAL float
#define PF_VAR 5
#include "pf_col.inc"
