// Library-side readers of the shape-tuning table (include/reform_tune.h).
#pragma once

namespace reform {

// Knobs of the batched PRIA step kernels ("batch.pria.*"); 0 = unset.
enum PriaKnob {
    kPriaXThreads = 0,   // FP32 x-tile workgroup: 256 or 512 threads
    kPriaXHold = 1,      // FP32 x tiles keep x1 in registers: 1 yes, 2 no
    kPriaYZShape = 2,    // y/z register shape: 1 compact where T' allows, 2 full
    kPriaKnobs = 3
};

// The table's value of a knob, 0 if unset.
int pria_knob(int which);

// Lanes per node of the batched and served tiers: the table's
// "batch.praq.lanes.<b>" for the mean degree, else the built-in rule.
int batch_praq_lanes(int n, int m);

}  // namespace reform
