// Layer table of the learned denoiser (net.hip).
#pragma once
#include <cstddef>

namespace lucida {

struct NetLayer {
    int cin, cout, relu;
};

constexpr int kNetNumLayers = 15;
constexpr int kNetInputsUsed = 15;  // feature channel 15 is zero
extern const NetLayer kNetLayers[kNetNumLayers];

// Parameters: per layer W[9][cin][cout] (tap = 3 * dy + dx) then b[cout], fp32.
size_t net_param_count();

}  // namespace lucida
