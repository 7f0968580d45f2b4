// solvers.hip -- device Krylov kernels (filled in below).
#include "internal.hpp"
