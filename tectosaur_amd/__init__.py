"""tectosaur_amd: the farfield / nearfield hot path of tectosaur as
hand-written HIP kernels for MI355X (gfx950) behind a C ABI
(include/tct_hip.h), with Python operators that keep tectosaur's operator
API (TriToTriDirectFarfieldOp, PairsIntegrator, RegularizedNearfieldIntegralOp,
RegularizedSparseIntegralOp)."""
