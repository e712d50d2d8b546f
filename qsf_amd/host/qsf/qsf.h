// qsf.h — umbrella header of the C++ host mirror (the reference's QSF.h plays this role).
#pragma once
#include "device_grid.h"
#include "propagator.h"
