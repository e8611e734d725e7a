// life_r3.hip -- Life band-walk kernels for Moore(3).
#define LIFE_R 3
#include "life_r.inc"
