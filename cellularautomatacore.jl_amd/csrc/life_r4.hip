// life_r4.hip -- Life band-walk kernels for Moore(4).
#define LIFE_R 4
#include "life_r.inc"
