#pragma once

// The device probe of an insertion: phase 1 of the wave builder with the
// layer-0 beams of the wave's level-0 nodes searched on the GPU (plain C++
// interface; the HIP side is build_probe.hip).

#include <cstdint>

#include "../host/graph.hpp"

namespace hnswamd {

/// What a device-probed add_batch did.
struct DeviceProbeStats {
  uint64_t waves = 0;        // waves probed on the device
  uint64_t device_nodes = 0; // nodes whose layer-0 beam ran on the GPU
  uint64_t host_nodes = 0;   // nodes probed on the host (upper-level nodes, early small waves)
  double device_ms = 0.0;    // device time of the probe launches (events)
  double refresh_ms = 0.0;   // host time spent validating and uploading changed lists
};

/// A ProbeFn for one add_batch. Waves below `min_device_wave` nodes before the
/// first device wave are probed on the host; from the first wave of at least
/// that size on, the graph lives on the GPU too and every later wave's
/// level-0 nodes are probed there, while the host probes the other nodes at
/// the same time. The result equals the host probe's exactly (the GPU search
/// computes the host's summation order and Alg. 2's order and rules).
/// `stats` (optional) must outlive the returned function's calls.
ProbeFn make_device_probe(uint32_t min_device_wave, DeviceProbeStats* stats);

}  // namespace hnswamd
