// Exposes the selected reference limiter macro (Src/Base/Slope_limiter/<name>/Slope_limiter.H)
// exactly as the schemes apply it (MUSCL_Scheme.cpp:83-85, :99-101): both arguments
// are written as unparenthesised differences. TEST INFRASTRUCTURE ONLY.
#include <cmath>
#include <cstdlib>
#include "Slope_limiter.H"
extern "C" double ref_lim_L(double u0, double u1, double u2) { return limiter(u1 - u0, u2 - u1); }
extern "C" double ref_lim_R(double u1, double u2, double u3) { return limiter(u3 - u2, u2 - u1); }
