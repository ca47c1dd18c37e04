#pragma once
#include "defines.hpp"
#include "lbm.hpp"
#include "shapes.hpp"
#include "units.hpp"
#include "info.hpp"

void main_setup(); // the scene, as in the reference (src/setup.hpp)
