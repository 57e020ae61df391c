// si_abi.hpp -- definitions of the opaque C ABI handles.
#pragma once

#include "si_host.hpp"
#include "../../include/spectral_mi355x.h"

struct si_context {
    si::Context c;
};

struct si_laplacian {
    si::DevLaplacian L;
    int device = 0;
};
