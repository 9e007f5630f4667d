// driver.hpp -- compatibility name for programs written against GPU-BSW's
// driver.hpp (reference driver.hpp:1-43): it declares the same
// gpu_bsw_driver API, implemented by libgbsw.so.
#ifndef GBSW_COMPAT_DRIVER_HPP
#define GBSW_COMPAT_DRIVER_HPP
#include <chrono>
#include <cmath>
#include <string>
#include <vector>

#include "gpu_bsw_driver.hpp"

#define NSTREAMS 2
#define NOW std::chrono::high_resolution_clock::now()
#endif
