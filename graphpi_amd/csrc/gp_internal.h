// Internal helpers shared by the host C++ sources and the HIP runtime.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "graphpi_amd.h"

// Records the calling thread's last error (read back by gp_last_error) and
// returns code, so call sites can write `return gp_fail(GP_EINVAL, "...")`.
int gp_fail(int code, const std::string& msg);

// Host CSR builder shared by the text loader, the in-memory edge path and the
// generators: 'pairs' holds directed (u,v) keys packed as u<<32|v.  Sorts,
// removes duplicates and fills the CSR; degree statistics come from 'degree'
// (counted before duplicate removal, as the text loader does).
int gp_build_csr(std::vector<uint64_t>& pairs, int32_t v_cnt, const std::vector<int64_t>& degree,
                 gp_csr** out);

// O(V+E) check that a CSR is safe to count on (offsets, id range, sorted
// lists without self-loops); GP_EINVAL with the reason otherwise.
int gp_csr_validate(const int32_t* edge, const uint32_t* vertex, int32_t v_cnt, uint32_t e_cnt);

// Parallel sort + unique of packed 64-bit keys.
void gp_sort_unique(std::vector<uint64_t>& keys);
