"""Every ensemble member's grid fields from one grouped batch.

The reference synthesizes the grid state of an ensemble one member at a time
(reference da/ensemble.py:547-579, ``grid_fields``: per member the model's
grid state, and ``inverse(inverse_laplacian(.))`` of vorticity and
divergence for the streamfunction and velocity potential), and steps the
members in a loop (da/ensemble.py:437-455).  ``EnsembleSynthesis`` stacks
the members' spectral fields and runs all of their syntheses -- winds,
potentials, theta, moisture, ln ps -- as one ``TransformBatch``: one pack, one
grouped synthesis GEMM and one set of inverse FFTs for the whole ensemble,
replayable as a hipGraph.  The inverse Laplacian is fused into the pack.
"""
from __future__ import annotations

from .batch import TransformBatch
from .transform import SphericalHarmonicTransform

__all__ = ["EnsembleSynthesis", "ensemble_grid_fields", "GRID_FIELDS", "SPECTRAL_SOURCES"]

#: grid fields the synthesis can produce, in the reference's names
GRID_FIELDS = ("u", "v", "theta", "qv", "lnps", "psi", "chi")
#: the member's spectral field each grid field is synthesized from
SPECTRAL_SOURCES = {"u": ("vorticity", "divergence"), "v": ("vorticity", "divergence"),
                    "theta": ("theta",), "qv": ("qv",), "lnps": ("log_surface_pressure",),
                    "psi": ("vorticity",), "chi": ("divergence",)}


def _member_field(member, name):
    """A spectral field of a member: a mapping, a state, or a state with an
    ``atmosphere`` (the reference's ``ArwenGlobalState``)."""
    if isinstance(member, dict):
        return member[name]
    holder = getattr(member, "atmosphere", member)
    return getattr(holder, name)


class EnsembleSynthesis:
    """Grid fields of ``members`` members with ``nlev`` levels each.

    ``load(members)`` copies the members' spectral fields into the stacked
    device inputs (or write them directly: ``inputs[name]`` has shape
    (members, nlev, T+1, T+1), (members, T+1, T+1) for
    ``log_surface_pressure``); ``run()`` returns ``{name: (members, ...)}``
    grid stacks, (members, nlev, nlat, nlon) or (members, nlat, nlon) for
    ``lnps``.  The outputs are the object's own buffers, overwritten by the
    next run.
    """

    def __init__(self, transform: SphericalHarmonicTransform, members: int, nlev: int,
                 names=GRID_FIELDS):
        names = tuple(names)
        unknown = [n for n in names if n not in GRID_FIELDS]
        if unknown:
            raise ValueError(f"unknown grid fields {unknown}; expected a subset of {GRID_FIELDS}")
        if members < 1 or nlev < 1:
            raise ValueError("an ensemble needs at least one member and one level")
        self.transform = transform
        self.members = int(members)
        self.nlev = int(nlev)
        self.names = names
        torch = transform.backend.xp
        dev = transform.backend.device()
        spec = transform.spectral_shape
        grid = transform.grid.shape
        R, L = self.members, self.nlev
        needed = sorted({s for n in names for s in SPECTRAL_SOURCES[n]})
        self.inputs = {
            s: torch.zeros((R, *spec) if s == "log_surface_pressure" else (R, L, *spec),
                           dtype=transform.backend.complex_dtype, device=dev)
            for s in needed}
        self.outputs = {
            n: torch.empty((R, *grid) if n == "lnps" else (R, L, *grid),
                           dtype=transform.backend.float_dtype, device=dev)
            for n in names}
        jobs = []
        x = self.inputs
        y = self.outputs
        if "u" in names or "v" in names:
            u = y.get("u")
            v = y.get("v")
            if u is None or v is None:  # the wind job writes both components
                u = u if u is not None else torch.empty_like(v)
                v = v if v is not None else torch.empty_like(u)
                self._spare = (u, v)
            jobs.append(("wind", transform, (x["vorticity"], x["divergence"]), (u, v)))
        for n in ("theta", "qv", "lnps"):
            if n in names:
                jobs.append(("inverse", transform, x[SPECTRAL_SOURCES[n][0]], y[n]))
        for n in ("psi", "chi"):
            if n in names:
                jobs.append(("inverse", transform, x[SPECTRAL_SOURCES[n][0]], y[n],
                             {"inverse_laplacian": True}))
        self.batch = TransformBatch(jobs)
        self._captured = False

    def load(self, members) -> None:
        """Copy each member's spectral fields into the stacked inputs."""
        members = list(members)
        if len(members) != self.members:
            raise ValueError(f"expected {self.members} members, got {len(members)}")
        for k, member in enumerate(members):
            for s, buf in self.inputs.items():
                src = self.transform._spectral_tensor(_member_field(member, s))
                if tuple(src.shape) != tuple(buf.shape[1:]):
                    raise ValueError(f"member {k} field {s} has shape {tuple(src.shape)}, "
                                     f"expected {tuple(buf.shape[1:])}")
                buf[k].copy_(src)

    def run(self, *, graph: bool = False) -> dict:
        """Synthesize every member's grid fields; ``graph`` replays the batch
        from a hipGraph (captured on the first such call)."""
        if graph:
            if not self._captured:
                self.batch.capture()
                self._captured = True
            self.batch.launch()
        else:
            self.batch.run()
        return self.outputs


def ensemble_grid_fields(transform, members, names=GRID_FIELDS, nlev=None) -> dict:
    """``{name: (R, ...)}`` grid stacks of every member, as the reference's
    ``grid_fields`` returns them (da/ensemble.py:547-579), from one batch."""
    members = list(members)
    if not members:
        raise ValueError("an ensemble needs at least one member")
    if nlev is None:
        probe = "vorticity" if any(n in names for n in ("u", "v", "psi")) else \
            "divergence" if "chi" in names else "theta" if "theta" in names else "qv"
        nlev = int(transform._spectral_tensor(_member_field(members[0], probe)).shape[0]) \
            if any(n != "lnps" for n in names) else 1
    ens = EnsembleSynthesis(transform, len(members), nlev, names)
    ens.load(members)
    return {n: t.clone() for n, t in ens.run().items()}
