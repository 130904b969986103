// Rotation kernel by (num_wann, nntot) on MI355X: the choices of the
// autotuner in w90gpu.hip (tune_member), measured over a sweep of synthetic
// problems by wannier90_amd/tune.py, which writes this file.  A key that no
// row covers keeps its class default.
// Source: rotation_tile_sweep.txt, 336 keys, 11 rows.
#pragma once

struct TuneRow {
  int nw_lo, nw_hi, nn_lo, nn_hi, rk;
};

constexpr TuneRow kTuneTable[] = {
    {33, 36, 1, 7, RK_DIRECT64},
    {49, 52, 1, 7, RK_DIRECT64},
    {65, 128, 1, 7, RK_GEMM32},
    {33, 36, 8, 10, RK_DIRECT64},
    {40, 40, 8, 10, RK_DIRECT64},
    {49, 52, 8, 10, RK_DIRECT64},
    {65, 128, 8, 10, RK_GEMM32},
    {33, 36, 11, 12, RK_DIRECT64},
    {38, 39, 11, 12, RK_DIRECT64},
    {49, 52, 11, 12, RK_DIRECT64},
    {65, 128, 11, 12, RK_GEMM32},
};
