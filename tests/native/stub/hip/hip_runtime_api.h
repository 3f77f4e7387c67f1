// Host-only stand-in: see hip_runtime.h beside it.
#pragma once
#include "hip_runtime.h"
