#pragma once
// The numa::hip C++20 layer: the reference's include/numa solver-path API
// (operators, providers, Krylov methods, Bratu3D, ChandrasekharOp) on the
// MI355X backend.  Code written against numa::op switches by namespace.
#include "hip/bratu3d.hpp"
#include "hip/chandrasekhar_op.hpp"
#include "hip/core.hpp"
#include "hip/gmres.hpp"
