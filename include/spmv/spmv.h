// Umbrella header of the C++ API (the reference's spmv/spmv.h).
#pragma once

#include "L2GMap.h"
#include "Matrix.h"
#include "Vector.h"
#include "cg.h"
#include "read_petsc.h"
