// dense.hpp -- the reference splits LinAlg into dense/vec/sparse headers; all three types live in
// linalg.hpp here, so either include path works.
#pragma once
#include "linalg.hpp"
