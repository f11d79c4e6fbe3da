"""tectosaur_amd: the farfield / nearfield hot path of tectosaur as
hand-written HIP kernels for MI355X (gfx950) behind a C ABI
(include/tct_hip.h), with Python operators that keep tectosaur's operator
API (tectosaur/__init__.py:17-44): TriToTriDirectFarfieldOp, FMMFarfieldOp,
PairsIntegrator, RegularizedNearfieldIntegralOp, RegularizedSparseIntegralOp,
RegularizedDenseIntegralOp, SumOp, MultOp, NegOp, CompositeOp, MassOp,
InteriorOp.

Importing the package loads nothing native; the operators load
libtct_hip.so on construction and raise when it or a GPU is missing."""
_EXPORTS = {
    'TriToTriDirectFarfieldOp': 'sparse_farfield_op',
    'FMMFarfieldOp': 'fmm',
    'PairsIntegrator': 'pairs_integrator',
    'RegularizedNearfieldIntegralOp': 'nearfield_op',
    'RegularizedSparseIntegralOp': 'sparse_integral_op',
    'RegularizedDenseIntegralOp': 'dense_integral_op',
    'SumOp': 'combinators',
    'MultOp': 'combinators',
    'NegOp': 'combinators',
    'CompositeOp': 'combinators',
    'DenseOp': 'combinators',
    'MassOp': 'mass_op',
    'InteriorOp': 'interior',
    'make_rect': 'mesh',
    'concat': 'mesh',
    'refine': 'mesh',
    'build_constraint_matrix': 'constraints',
    'continuity_constraints': 'constraints',
    'free_edge_constraints': 'constraints',
    'find_free_edges': 'constraints',
    'free_edge_dofs': 'constraints',
}


def __getattr__(name):
    mod = _EXPORTS.get(name)
    if mod is None:
        raise AttributeError("module 'tectosaur_amd' has no attribute %r" % name)
    import importlib
    return getattr(importlib.import_module('.' + mod, __name__), name)


def __dir__():
    return sorted(list(globals()) + list(_EXPORTS))
