// life_r1.hip -- Life band-walk kernels for Moore(1).
#define LIFE_R 1
#include "life_r.inc"
