"""smolyax_amd: barycentric Smolyak interpolation evaluated by hand-written HIP kernels on MI355X.

Drop-in for the evaluation API of JoWestermann/smolyax:
    from smolyax_amd import nodes, indices
    from smolyax_amd.interpolation import SmolyakBarycentricInterpolator
"""

from . import nodes  # noqa: F401

__version__ = "0.1.0"


def __getattr__(name):
    # native-backed modules load lazily so that `import smolyax_amd.nodes` works without the library
    if name in ("indices", "interpolation", "barycentric", "packing"):
        import importlib

        return importlib.import_module(f".{name}", __name__)
    raise AttributeError(name)
