"""The C ABI without a GPU: the library loads, exports every symbol the header declares,
its ctypes mirror has the header's struct layouts, and a context cannot be created
without a device (no CPU fallback)."""
import ctypes as C
import os
import re
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "particulator_hip.h")


def declared_functions():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\**\s*(pt_[a-z_0-9]+)\s*\(", txt, re.M)))


def test_header_declares_the_step():
    names = declared_functions()
    for n in ("pt_create", "pt_destroy", "pt_upload_table", "pt_upload_sb", "pt_upload_pe",
              "pt_set_field", "pt_add_particles", "pt_init", "pt_advance", "pt_run", "pt_compact",
              "pt_stats", "pt_download", "pt_last_error"):
        assert n in names


def test_library_exports_every_declared_symbol():
    from particulator_hip._lib import LIBPATH, SIGNATURES
    assert os.path.exists(LIBPATH), "build the library first (__graft_entry__.build())"
    lib = C.CDLL(LIBPATH)
    for n in declared_functions():
        assert hasattr(lib, n), n
        assert n in SIGNATURES, "ctypes binding lacks %s" % n
    assert lib.pt_abi_version() == 1


def test_ctypes_layouts_match_the_header():
    from particulator_hip import _lib as L
    structs = {"pt_proc_desc": L.ProcDesc, "pt_table_desc": L.TableDesc, "pt_field_fn": L.FieldFn,
               "pt_field_desc": L.FieldDesc, "pt_config": L.Config, "pt_pop_stats": L.PopStats,
               "pt_work": L.Work, "pt_process": L.Process, "pt_launch": L.Launch,
               "pt_wall_record": L.WallRecord, "pt_tune_trial": L.TuneTrial, "pt_learn_config": L.LearnConfig,
               "pt_learn_report": L.LearnReport}
    src = ["#include <stdio.h>", "#include <stddef.h>", '#include "%s"' % HEADER, "int main(void){"]
    for cname, cls in structs.items():
        src.append('printf("%s %%zu\\n", sizeof(%s));' % (cname, cname))
        for f, _ in cls._fields_:
            src.append('printf("%s.%s %%zu\\n", offsetof(%s, %s));' % (cname, f, cname, f))
    src.append("return 0;}")
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "t.c")
        open(c, "w").write("\n".join(src))
        exe = os.path.join(d, "t")
        subprocess.run(["gcc", "-std=c99", "-o", exe, c], check=True)
        out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split("\n")
    got = dict(line.split() for line in out if line.strip())
    for cname, cls in structs.items():
        assert int(got[cname]) == C.sizeof(cls), cname
        for f, _ in cls._fields_:
            assert int(got["%s.%s" % (cname, f)]) == getattr(cls, f).offset, (cname, f)


def test_no_device_means_no_context():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    import particulator_hip as P
    with pytest.raises(P.PTError) as e:
        P.MultiPopulation(seed=0, electron=P.Population(16, [], None, 0.0, species=0))
    assert e.value.code in (-9, -2)
