// knc_mlp_host.hpp -- host-side state of a learned-guess network: FP32
// master parameters, the normalisation constants and the device copy of the
// MFMA operand fragments used by inference.
#pragma once

#include <vector>

#include "knc_mlp.hpp"

namespace knc {

struct MlpHost {
    int device = 0;
    std::vector<float> params;   // MLP_NPARAM
    std::vector<float> norm;     // MLP_NNORM
    void* d_frag = nullptr;
    MlpDev dev{};
    int upload_fragments();
    static size_t frag_bytes();
};

}  // namespace knc
