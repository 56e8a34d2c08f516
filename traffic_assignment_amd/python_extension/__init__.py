"""Python operators over the gfx950 engine: the reference's `pape` module
interface (pape.py) and the differentiable equilibrium (autograd.py)."""
