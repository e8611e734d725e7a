// life_r2.hip -- Life band-walk kernels for Moore(2).
#define LIFE_R 2
#include "life_r.inc"
