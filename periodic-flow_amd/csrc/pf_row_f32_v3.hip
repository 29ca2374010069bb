#define PF_REAL float
#define PF_VAR 3
#include "pf_row.inc"
