"""The reference program's run loop (src/main.c:348-611) for Python callers:
the config.h settings of a run as arguments, the physics parameters as
keywords (the defaults are config.h's), one call to sf_run."""
from __future__ import annotations

import ctypes as C

from . import _lib as L


def run(num_bodies: int, num_timesteps: int, *, output_interval: int = 0, enable_sim: bool = True,
        enable_vis: bool = False, sim_filename: str | None = "sim/step_%04u", vis_filename: str | None = None,
        vis_file_type: int = L.IMAGE_FILETYPE_PNG, vis_size: tuple[int, int] = (1080, 1080),
        init_cond_file: str | None = None, verbose: bool = False, **params) -> int:
    """Runs steps 0 .. num_timesteps under params["out_dir"] (OUTPUT_DIR).
    Returns 0 when the run ends or a stop file ends it, 1 when it ends early
    because a replayed snapshot could not be read; raises when the set-up
    fails."""
    lib = L.load()
    L.set_params(**params)
    enc = (lambda s: s.encode() if s is not None else None)
    cfg = L.RunConfig(num_bodies=num_bodies, num_timesteps=num_timesteps, output_interval=output_interval,
                      enable_sim=int(enable_sim), enable_vis=int(enable_vis), sim_filename=enc(sim_filename),
                      vis_filename=enc(vis_filename), vis_file_type=vis_file_type, vis_sizex=vis_size[0],
                      vis_sizey=vis_size[1], init_cond_file=enc(init_cond_file), verbose=int(verbose))
    rc = lib.sf_run(C.byref(cfg))
    if rc < 0:
        raise RuntimeError(f"sf_run failed: {L.last_error()}")
    return rc
