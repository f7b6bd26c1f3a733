// Umbrella header under the reference's name (KittlesPT/include/kittlesPT/kittlesPT.hpp:1-2).
#pragma once
#include "renderer.hpp"
