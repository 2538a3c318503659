#pragma once
#include <vector>

#include "../../include/jpxlg.h"
#include "internal.h"

namespace jpxlg {

void build_const_tables(ConstTables* t);
// The L matrices of the matrix-core DCT forms (f64 closed form, rounded once).
void build_dct_mats(DctMats* m);
void srgb8_lut(float* lut256);
void srgb16_lut(int bits, float* lut65536);
int depth_tables(int bits, std::vector<float>* thresholds, std::vector<float>* lut);
void build_probe_params(const jpxlg_quantizer& q, int sharpness, ProbeParams* p);
// The cover objective of a structure build: legacy = the fixed-quantizer
// request's (legacy rate proxy, flat cell weights), else the calibrated rate
// model with quantizer-donor weights.
void set_cover_objective(ProbeParams* p, bool legacy);
// I.3.2 natural coefficient order of a square DCT (position -> cell).
void natural_order_table(int side, uint16_t* order);
float qm_multiplier(uint32_t qm_scale);
float hf_multiplier(uint32_t global_scale, uint32_t hf_mul);

}  // namespace jpxlg
