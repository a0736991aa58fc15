"""Write a synthetic workload as reference-format input files.

    python -m wavesim.casefiles <case> <directory> [NT]

<case> is one of the generators in wavesim.cases (acoustic_c1, elastic_c2,
tmem_c3, elastic_c4, elastic_c5).  Writes <directory>/config.txt (the
reference's KEY=VALUE configuration, Configuration.cpp:45-87), the model as
MatrixMarket vectors (<directory>/model.{vp,vs,density}.mtx or
{epsilonr,sigma,mur}.mtx, IO.hpp:21-164), the source file (one Ricker per
shot, AcquisitionSettings.hpp:55-100) and the receiver file(s) (one shared file,
or receivers.shot_<n>.txt when shots differ: Receivers.cpp:226-248), so that
`bin/wavesim_run <directory>/config.txt` runs the same workload.
"""
from __future__ import annotations

import os
import sys

import numpy as np

from . import cases


def _vec(path, v):
    with open(path, "w") as f:
        f.write("%%MatrixMarket matrix array real general\n%d 1\n" % len(v))
        np.savetxt(f, np.asarray(v, np.float64), fmt="%.8g")


def write_case(c, out, fc=None):
    os.makedirs(out, exist_ok=True)
    s = c.settings
    NX, NY, NZ = s["NX"], s["NY"], s.get("NZ", 1)
    em = c.type == "tmem"
    names = ("epsilonr", "sigma", "mur") if em else ("vp", "vs", "density")
    for name, v in zip(names, (c.m0, c.m1, c.m2)):
        if v is not None:
            _vec(os.path.join(out, f"model.{name}.mtx"), v)

    def xyz(i):
        i = int(i)
        return i % NX, i // (NX * NZ), (i // NX) % NZ

    freq = fc if fc is not None else {"tmem": 2e8}.get(c.type, 30.0)
    with open(os.path.join(out, "sources.txt"), "w") as f:
        for k, sh in enumerate(c.shots):
            for i, t in zip(sh.src_index, sh.src_type):
                x, y, z = xyz(i)
                f.write(f"{k + 1} {x} {y} {z} {int(t)} 1 1 {freq} 1.0 0.0\n")
    same = all(np.array_equal(sh.rec_index, c.shots[0].rec_index) and np.array_equal(sh.rec_type, c.shots[0].rec_type)
               for sh in c.shots)
    files = [("receivers.txt", c.shots[0])] if same else [(f"receivers.shot_{k + 1}.txt", sh)
                                                          for k, sh in enumerate(c.shots)]
    for name, sh in files:
        with open(os.path.join(out, name), "w") as f:
            for i, t in zip(sh.rec_index, sh.rec_type):
                x, y, z = xyz(i)
                f.write(f"{x} {y} {z} {int(t)}\n")
    DT = s["DT"]
    lines = [f"dimension={c.dimension}", f"equationType={c.type}", f"NX={NX}", f"NY={NY}", f"NZ={NZ}",
             f"DH={s['DH']}", f"DT={DT!r}", f"T={s['NT'] * DT!r}", f"spatialFDorder={s['spatial_fd_order']}",
             "ModelRead=1", f"ModelFilename={os.path.join(out, 'model')}", "fileFormat=1",
             f"FreeSurface={s.get('free_surface', 0)}", f"DampingBoundary={s.get('damping_boundary', 0)}",
             f"BoundaryWidth={s.get('boundary_width', 0)}", f"DampingCoeff={s.get('damping_coeff', 8.0)}",
             f"VMaxCPML={s.get('vmax_cpml', 3500.0)}", f"CenterFrequencyCPML={s.get('fc_cpml', 5.0)}",
             f"NPower={s.get('npower', 4.0)}", f"SourceFilename={os.path.join(out, 'sources')}",
             f"ReceiverFilename={os.path.join(out, 'receivers')}", f"useReceiversPerShot={0 if same else 1}",
             f"SeismogramFilename={os.path.join(out, 'seis')}", f"seismoDT={DT!r}"]
    with open(os.path.join(out, "config.txt"), "w") as f:
        f.write("\n".join(lines) + "\n")
    return os.path.join(out, "config.txt")


def main(argv):
    if len(argv) < 3:
        print(__doc__)
        return 2
    gen = getattr(cases, argv[1])
    kw = {"NT": int(argv[3])} if len(argv) > 3 else {}
    print(write_case(gen(**kw), argv[2]))
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv))
