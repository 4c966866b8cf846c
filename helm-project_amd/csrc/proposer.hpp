// proposer.hpp -- learned candidate proposal for the conservative broad phase
// (proposer.hip) and its statistics.
#pragma once

#include "axis_hip.hpp"
#include "mlp.hpp"

namespace axis_hip {

/// What the learned proposer did in one generation.
struct ProposerStats {
    index_t rows = 0;            // destination cells
    index_t guard_rejected = 0;  // proposals that failed the containment guard -> LBVH
    index_t area_failed = 0;     // accepted proposals whose overlaps missed the cell area by > 1e-12 -> LBVH
    index_t pairs_proposed = 0;  // candidate pairs of the accepted proposals
    index_t pairs_fallback = 0;  // candidate pairs the LBVH returned for the redone cells
    double gen_ms = 0.0;
};

namespace detail {

/// Candidates of every destination cell from the network's proposals:
/// cand_off [n_dst + 1], cand_src, accepted[j] = 1 where the proposal passed
/// the guard (rejected rows have no candidates).
void proposer_candidates(const MeshGeom &sg, const MeshGeom &dg, const mlp::Model &model, DeviceBuffer<index_t> &cand_off,
                         DeviceBuffer<std::int32_t> &cand_src, DeviceBuffer<std::int32_t> &accepted, hipStream_t s);

/// Training pairs for the network: x [n_dst][6] features, t [n_dst][4]
/// labels (zero centre offset, log of the guard radius / feature_scale), device pointers.
void proposer_training_data(const Mesh &src, const Mesh &dst, float feature_scale, float *x, float *t, hipStream_t s, index_t *n_unlabelled);

}  // namespace detail
}  // namespace axis_hip
