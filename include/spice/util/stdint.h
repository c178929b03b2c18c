#pragma once
// Integer aliases used throughout the public API. The reference spells its API with these
// names (spice/util/stdint.h:7-13), so a drop-in keeps them.
#include <cstddef>
#include <cstdint>

using int_ = std::int32_t;
using uint_ = std::uint32_t;
using long_ = std::int64_t;
using ulong_ = std::uint64_t;
using size_ = std::size_t;
