// Reverse-mode (adjoint) arithmetic of fvm_math.hpp.
#pragma once
#include "fvm_math.hpp"
