// k_adjoint.hip -- reverse-mode kernels of the collisionless step (filled in below).
#include "launch.h"
