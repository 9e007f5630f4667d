// utils.hpp -- compatibility name for GPU-BSW's utils.hpp (reference
// utils.hpp:23-25): free_alignments() and getMaxLength().
#ifndef GBSW_COMPAT_UTILS_HPP
#define GBSW_COMPAT_UTILS_HPP
#include <string>
#include <vector>

#include "gpu_bsw_driver.hpp"

inline unsigned getMaxLength(std::vector<std::string> v)
{
    unsigned m = 0;
    for (const auto& s : v) m = s.length() > m ? (unsigned)s.length() : m;
    return m;
}
#endif
