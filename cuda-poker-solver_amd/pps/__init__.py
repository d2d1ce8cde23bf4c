"""MI355X postflop solver (DCFR over a compiled bet tree, gfx950 HIP kernels,
hipGraph replay) — drop-in for the reference's `pps` package
(python/pps/__init__.py).

    import pps
    s = pps.Solver(board="Qs9h2d", oop="TT+,AKo", ip="AQ+,KQs",
                   pot=200, stack=500, bets="0.75,a", raises="2.5,3")
    s.solve(2000)            # or s.solve(max_iters=100000, target=0.1)
    s.stats()                # {"ev_oop", "ev_ip", "exploitability", "pair_mass"}
    s.strategy()             # labeled per-combo root strategy
    s.save("spot.sol"); t = pps.Solver.load("spot.sol"); t.continue_solve(500)

Serving many spots (no reference counterpart in the module: the batching
layer under the HTTP API, pps.api):

    srv = pps.Server()                       # one device arena
    a = srv.add(board="Qs9h2d7c", oop="TT+,AKo", ip="AQ+,KQs", pot=200, stack=500)
    a.solve(2000)        # from several threads at once: solves run batched
    a.start_solve(2000); a.job(); a.wait()   # or as a background job

The native extension is built in-tree by `make -C cuda-poker-solver_amd`
(or __graft_entry__.build()). There is no CPU fallback: importing fails
loudly when the extension is missing.
"""

__all__ = ["Solver", "Server", "ArenaFull", "__version__"]

__version__ = "0.1.0"

try:
    from . import pps_native as _native  # noqa: F401
except ImportError as e:  # pragma: no cover - build not run
    raise ImportError(
        "the pps native module is not built; run `make -C cuda-poker-solver_amd` "
        "(hipcc for gfx950 + pybind11)") from e

Solver = _native.Solver
Server = _native.Server        # many spots on one GPU, solves batched
ArenaFull = _native.ArenaFull  # Server.add / load: no room in the device arena
