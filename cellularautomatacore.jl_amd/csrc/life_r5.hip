// life_r5.hip -- Life band-walk kernels for Moore(5).
#define LIFE_R 5
#include "life_r.inc"
