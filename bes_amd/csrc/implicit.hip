// Implicit path of the BES solver core (filled in below).
#include <hip/hip_runtime.h>

#include "../../include/bes_hip.h"
#include "context.hpp"

extern "C" void bes_implicit_free(bes_ctx *c) { (void)c; }
