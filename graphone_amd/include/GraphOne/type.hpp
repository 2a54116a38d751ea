// Core types of the graph_one C++ API (the reference's include/GraphOne/type.hpp:8-25).
//
// arch_t names where a buffer lives: arch_t::cpu (host memory) or arch_t::hip (HBM of the MI355X the
// process runs on). Code written for the reference names the device arch_t::cuda; that one enumerator
// is the documented source change of a drop-in port (see the repository's integration notes).
#pragma once

#include <cstdint>
#include <string>

namespace graph_one {

using vid_t = uint32_t;
using eid_t = uint32_t;

enum class arch_t {
    cpu,
    hip
};

struct empty_t {
    operator std::string() const { return ""; }
};

enum class vstart_t {
    FROM_0_TO_0,
    FROM_1_TO_0,
};

}  // namespace graph_one
