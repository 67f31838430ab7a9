// Library-internal accessors shared between the pipeline (pipeline.hip)
// and the C++ drop-in layer (dropin.cpp); not part of the C ABI.
#pragma once

struct ops_workspace;

namespace opsamd {

// The workspace's list of row ids grouped by numeric bin after its last
// call (device memory, valid until the next call; null before the first).
const int *workspace_bins(const ops_workspace *ws);

}  // namespace opsamd
