// life_r6.hip -- Life band-walk kernels for Moore(6).
#define LIFE_R 6
#include "life_r.inc"
