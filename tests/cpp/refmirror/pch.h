// The reference compiles every source with its precompiled header, which brings
// BlazeEngine in with `using namespace Blaze` (SPHFluidSimulation/source/pch.h);
// this is the same for the test build.
#pragma once
#include "blaze_min.h"

using namespace Blaze;
