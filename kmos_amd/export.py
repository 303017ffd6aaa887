"""Export backend: write a model directory the engine runs from.

The counterpart of ``kmos.io.export_source(project, export_dir,
code_generator)`` (kmos/io/__init__.py:3884-3974).  Where kmos writes
base.f90 / lattice.f90 / proclist.f90 plus kmc_settings.py and builds them
into the f2py module ``kmc_model``, this backend writes

  kmc_settings.py   same keys as kmos's (io:3725-3845): model_name,
                    simulation_size, random_seed, setup_model, parameters,
                    rate_constants {name: (expr, enabled)}, site_names,
                    representations, lattice_representation, species_tags,
                    tof_count, xml, plus ``backend``
  kmcg_model.npy    the compiled op program (int32 blob, include/kmcg.h)
  kmc_model.py      three lines binding ``base``, ``lattice``, ``proclist``
                    (and for otf ``proclist_pars``) to the engine
                    (kmos_amd.kmc_model)

so that ``from kmc_model import base, lattice, proclist`` in the export
directory gives the f2py surface KMC_Model uses, running on the GPU.

Command line:  python -m kmos_amd.export MODEL.xml|MODEL.ini EXPORT_DIR [-b local_smart|lat_int|otf]
"""
from __future__ import annotations

import argparse
import os
import sys

import numpy as np

from .compiler import compile_model
from .model import Project, load_project, to_xml

KMC_MODEL_STUB = '''"""f2py-compatible kmc_model module of this export, run by kmos_amd."""
from kmos_amd.kmc_model import module_for_export as _m
base, lattice, proclist, proclist_pars = _m(__file__)
'''


def settings_source(pt: Project, backend: str) -> str:
    site_names = ["%s_%s" % ls for ls in pt.site_list()]
    lines = [
        "model_name = %r" % pt.model_name,
        "simulation_size = 20",
        "random_seed = 1",
        "backend = %r" % backend,
        "",
        "def setup_model(model):",
        '    """Write initialization steps here."""',
        "    pass",
        "",
        "# Default history length in graph",
        "hist_length = 30",
        "",
        "parameters = {",
    ]
    for name, par in pt.parameters.items():
        lines.append('    "%s":{"value":"%s", "adjustable":%s, "min":"%s", "max":"%s","scale":"%s"},'
                     % (name, par.get("value"), bool(par.get("adjustable", False)), par.get("min", 0.0),
                        par.get("max", 0.0), par.get("scale", "linear")))
    lines += ["    }", "", "rate_constants = {"]
    for p in pt.processes:
        lines.append('    "%s":("%s", %s),' % (p.name, p.rate_constant, bool(p.enabled)))
    lines += ["    }", "", "site_names = %r" % site_names, "representations = {"]
    for sp in sorted(pt.species):
        lines.append('    "%s":"""%s""",' % (sp, pt.representations.get(sp, "").strip()))
    lines += ["    }", "", 'lattice_representation = """"""', "", "species_tags = {"]
    for sp in sorted(pt.species):
        lines.append('    "%s":"""""",' % sp)
    lines += ["    }", "", "tof_count = {"]
    for p in sorted(pt.processes, key=lambda q: q.name):
        if p.tof_count is not None:
            lines.append('    "%s":%r,' % (p.name, p.tof_count))
    lines += ["    }", "", 'xml = """%s"""' % to_xml(pt), ""]
    return "\n".join(lines)


def export_source(project, export_dir: str, code_generator: str = "local_smart") -> str:
    """Write the export directory for ``project`` (a Project or a model file
    path) and ``code_generator`` ('local_smart', 'lat_int' or 'otf')."""
    pt = load_project(project) if isinstance(project, str) else project
    cm = compile_model(pt, code_generator)
    os.makedirs(export_dir, exist_ok=True)
    np.save(os.path.join(export_dir, "kmcg_model.npy"), cm.to_blob())
    with open(os.path.join(export_dir, "kmc_settings.py"), "w") as fh:
        fh.write(settings_source(pt, code_generator))
    with open(os.path.join(export_dir, "kmc_model.py"), "w") as fh:
        fh.write(KMC_MODEL_STUB)
    return export_dir


def main(argv=None):
    ap = argparse.ArgumentParser(description="export a kmos model for the MI355X engine")
    ap.add_argument("model")
    ap.add_argument("export_dir")
    ap.add_argument("-b", "--backend", default="local_smart", choices=["local_smart", "lat_int", "otf"])
    a = ap.parse_args(argv)
    print(export_source(a.model, a.export_dir, a.backend))
    return 0


if __name__ == "__main__":
    sys.exit(main())
