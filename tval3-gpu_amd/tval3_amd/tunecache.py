"""A small JSON cache of the launch shapes timed on this GPU.

The fused solver's layout (`Operator.prepare`, tv3_op_prepare) and a batched
solve (`api.tval3_gpu_3d_batched`, tv3_solve_batched) time their launch
shapes the first time they meet an operator and volume shape: 3-4 s at 128^3,
5-6 s for the config-3 batch.  `TuneCache` keeps the shapes they chose in a
JSON file, keyed by the GPU and the problem's sizes, and replays them
(tv3_op_layout_force builds a layout it is handed shapes for without timing
anything; a batch takes its shapes through the batch info's in-fields), so a
later process skips those timings.

A key holds sizes, not matrix contents: operators of the same kind, sizes,
entry count and weight type share an entry.  Any launch shape gives the same
results up to the FP32 summation order, so a stale entry costs speed, not
correctness.  A file that cannot be read, or one of another format version,
is ignored and rewritten; an entry the library refuses is timed again and
replaced.

    cache = TuneCache("tune.json")
    info = cache.prepare(A, dims)                  # layout, shapes timed once
    Us, infos, binfo, _ = cache.batched(ctx, ops, bs, dims_list, opts)
"""
import json
import os
import tempfile

VERSION = 1


def _device_key(ctx):
    import torch
    p = torch.cuda.get_device_properties(ctx.device)
    return f"{p.name}|{getattr(p, 'gcnArchName', '')}"


def _op_key(op, dims=None):
    k = f"{op.kind}:{op.m}x{op.n}:{op.nnz}:{getattr(op, 'dtype', '')}"
    return k if dims is None else k + ":" + "x".join(str(int(v)) for v in dims)


class TuneCache:
    def __init__(self, path, device=None):
        """path: the JSON file (created on the first store); device: the GPU's
        key (default: its name and architecture as torch reports them)."""
        self.path = path
        self.device = device
        self.entries = {}
        try:
            with open(path) as f:
                d = json.load(f)
            if (isinstance(d, dict) and d.get("version") == VERSION
                    and isinstance(d.get("entries"), dict)):
                self.entries = d["entries"]
        except (OSError, ValueError):
            pass

    def save(self):
        """Write the file atomically (a temporary file renamed over it)."""
        folder = os.path.dirname(os.path.abspath(self.path))
        fd, tmp = tempfile.mkstemp(dir=folder, prefix=".tunecache-", suffix=".json")
        try:
            with os.fdopen(fd, "w") as f:
                json.dump({"version": VERSION, "entries": self.entries}, f, indent=1,
                          sort_keys=True)
            os.replace(tmp, self.path)
        except BaseException:
            if os.path.exists(tmp):
                os.remove(tmp)
            raise

    def _dev(self, ctx):
        return self.device if self.device is not None else _device_key(ctx)

    def layout_key(self, op, dims):
        return "layout|" + self._dev(op.ctx) + "|" + _op_key(op, dims)

    def prepare(self, op, dims):
        """op.prepare(dims) with the shapes of an earlier run replayed when the
        cache has them (no timing), else timed and stored.  Returns the
        layout description."""
        from ._lib import TV3Error
        if op.kind == "dynamic":     # a matrix-free layout has no launch shapes to time
            return op.prepare(dims)
        key = self.layout_key(op, dims)
        e = self.entries.get(key)
        if e is not None:
            try:
                return op.layout_force(dims, e["code_N"], e["code_T"], e["pan_nnz"],
                                       e["panels"])
            except (TV3Error, KeyError, TypeError):
                pass
        info = op.prepare(dims)
        self.entries[key] = dict(code_N=int(info["code_N"]), code_T=int(info["code_T"]),
                                 pan_nnz=16384, panels=int(info["panels"]))
        self.save()
        return info

    def batch_key(self, ctx, ops, dims):
        # problems sharing one operator object (a dense group) name its first index
        first = {}
        parts = []
        for k, (op, d) in enumerate(zip(ops, dims)):
            j = first.setdefault(id(op), k)
            parts.append(_op_key(op, d) + (f"@{j}" if j != k else ""))
        return "batch|" + self._dev(ctx) + "|" + ";".join(parts)

    def batched(self, ctx, ops, bs, dims, opts=None, **kw):
        """api.tval3_gpu_3d_batched with the launch shapes of an earlier run of
        the same batch replayed (no timing), else timed and stored.  Shapes
        given in kw take precedence."""
        from . import api
        from ._lib import TV3Error
        key = self.batch_key(ctx, ops, dims)
        e = self.entries.get(key)
        if e is not None:
            shapes = dict(lanes=tuple(e["lanes"]), panels=e["panels"], panels_A=e["panels_A"],
                          panel_box_A=e["panel_box_A"])
            if e.get("splits"):
                shapes["splits"] = tuple(e["splits"])
            try:
                return api.tval3_gpu_3d_batched(ctx, ops, bs, dims, opts, **{**shapes, **kw})
            except (TV3Error, KeyError, TypeError):
                pass
        out = api.tval3_gpu_3d_batched(ctx, ops, bs, dims, opts, **kw)
        bi = out[2]
        self.entries[key] = dict(
            lanes=[int(bi["lanes_N"]), int(bi["lanes_T"])],
            splits=[int(bi["splits_N"]), int(bi["splits_T"])] if bi["dense_groups"] else None,
            panels=int(bi["panels_T"]) or -1, panels_A=int(bi["panels_A"]) or -1,
            panel_box_A=int(bi["panel_box_A"]))
        self.save()
        return out
