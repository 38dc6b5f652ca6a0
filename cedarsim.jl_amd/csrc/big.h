// big.h — the large-circuit path (one circuit of 10^4..10^6 unknowns): the
// supernodal multifrontal LU of bigplan.h on the device, with device-wide
// evaluation, assembly and a device-resident step controller.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/cedarhip.h"
#include "bigplan.h"

namespace cedar {

struct BigCircuit;

// Builds the plan, uploads it and allocates the workspace on the current device.
int big_create(const cs_netlist* nl, hipStream_t s, BigCircuit** out, std::string& err);
void big_destroy(BigCircuit* b);
const BigPlan& big_plan(const BigCircuit* b);
void big_info(const BigPlan& B, cs_big_info* info);

// Transient and DC operating point with the semantics of cs_tran / cs_dcop.
int big_tran(BigCircuit* b, hipStream_t s, const double* p, const double* u0, const cs_tran_opts* o,
             double* wave, double* u_final, cs_tran_stats* st, std::string& err);
int big_dcop(BigCircuit* b, hipStream_t s, const double* p, const cs_tran_opts* o, double* u_out,
             cs_tran_stats* st, std::string& err);

// DAE interface on the natural pattern (cs_residual / cs_jacobian) and the
// factor / solve pair (cs_lu_refactor / cs_lu_solve) on natural CSR values.
int big_residual(BigCircuit* b, hipStream_t s, double* out, const double* du, const double* u,
                 const double* p, double t, int32_t mode, std::string& err);
int big_jacobian(BigCircuit* b, hipStream_t s, double* vals, const double* du, const double* u,
                 const double* p, double gamma, double t, int32_t mode, std::string& err);
int big_lu_refactor(BigCircuit* b, hipStream_t s, const double* vals, std::string& err);
int big_lu_solve(BigCircuit* b, hipStream_t s, double* x, std::string& err);

// Unknowns recorded in waveforms (natural indices, validated by the caller).
int big_set_probes(BigCircuit* b, const std::vector<int32_t>& idx, std::string& err);

// Per-launch timing of one factorization and one solve (cs_big_profile).
int big_profile(BigCircuit* b, hipStream_t s, int32_t* kinds, int32_t* levels, float* ms, int32_t cap,
                int32_t* count, std::string& err);

// Time of the last factorization and of the last forward+backward solve
// (HIP events on the stream), milliseconds.
void big_last_times(const BigCircuit* b, double* factor_ms, double* solve_ms);

}  // namespace cedar
