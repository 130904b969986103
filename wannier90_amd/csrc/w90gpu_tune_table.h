// Rotation kernel by (num_wann, nntot) on MI355X: the choices of the
// autotuner in w90gpu.hip (tune_member), measured over a sweep of synthetic
// problems by wannier90_amd/tune.py, which writes this file.  A key that no
// row covers keeps its class default.
// Source: rotation_tile_sweep.txt, 336 keys, 9 rows.
#pragma once

struct TuneRow {
  int nw_lo, nw_hi, nn_lo, nn_hi, rk;
};

constexpr TuneRow kTuneTable[] = {
    {33, 48, 1, 7, RK_FUSED48},
    {49, 52, 1, 7, RK_DIRECT64},
    {65, 128, 1, 7, RK_GEMM32},
    {33, 48, 8, 10, RK_FUSED48},
    {49, 52, 8, 10, RK_DIRECT64},
    {65, 128, 8, 10, RK_GEMM32},
    {33, 48, 11, 12, RK_FUSED48},
    {49, 52, 11, 12, RK_DIRECT64},
    {65, 128, 11, 12, RK_GEMM32},
};
