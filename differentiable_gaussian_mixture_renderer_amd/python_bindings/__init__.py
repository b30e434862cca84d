"""Drop-in Python entry points: `vol_marcher.VolMarcher` / `vol_marcher.apply`
(reference: python_bindings/vol_marcher.py) and `splatter.Splatter` /
`splatter.apply` with the same argument list."""
