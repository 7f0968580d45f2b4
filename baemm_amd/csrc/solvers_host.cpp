// solvers_host.cpp -- host drivers of the Krylov solvers.
#include "internal.hpp"
