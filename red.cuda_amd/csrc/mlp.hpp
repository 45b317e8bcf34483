// mlp.hpp -- internal interface between the integrator and the step proposer.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

struct rd_mlp;

namespace rdamd {
// Device BF16 weights of a proposer.
const uint16_t* mlp_weights(const rd_mlp* p);
// Device ordinal the proposer's buffers live on.
int mlp_device(const rd_mlp* p);
// y[k] = log h proposal for raw feature rows x[k] (device pointers), on stream st.
int mlp_forward_dev(const rd_mlp* p, hipStream_t st, const float* x, uint32_t n, float* y);
// Sets the thread's rd_last_error() text.
void set_last_error(const std::string& msg);
}  // namespace rdamd
