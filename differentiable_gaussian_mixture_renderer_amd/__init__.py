"""dgmr for MI355X: the tile splatter and volumetric ray marcher of the
differentiable Gaussian mixture renderer as hand-written gfx950 kernels behind
a C ABI (include/dgmr.h), with the reference's Python operator interface in
`python_bindings`."""
__version__ = "0.1.0"
