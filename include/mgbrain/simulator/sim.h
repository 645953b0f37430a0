// sim.h -- simulator include path of the reference (src/simulator/sim.h:1).
#pragma once
#include "mgsim.h"
