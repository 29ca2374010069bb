#define PF_REAL float
#define PF_VAR 0
#include "pf_row.inc"
