// psc_model.hpp -- the potential of this build: the model header named by
// PSC_MODEL_HEADER (default models/preheat.hpp), chosen at build time as the
// reference chooses its model.h.  Included by the device code, the host
// initial-condition generator and the CPU oracle alike.
#pragma once
#include "psc_model_api.hpp"

#ifndef PSC_MODEL_HEADER
#define PSC_MODEL_HEADER "models/preheat.hpp"
#endif
#include PSC_MODEL_HEADER

namespace psc {
using Model = PSC_MODEL;
}  // namespace psc
