"""FNV-1a 64, the hash of the checkpoint payload, manifest, scenario and
decomposition (solver.rs:831-850), pinned by its published test vectors
(Fowler/Noll/Vo): the library's and the host module's implementations."""
import ctypes as C

import lbmflow_amd as L
from lbmflow_amd import checkpoint as K

VECTORS = [(b"", 0xCBF29CE484222325), (b"a", 0xAF63DC4C8601EC8C), (b"foobar", 0x85944171F73967E8)]


def test_fnv1a64_published_vectors():
    lib = L.lib()
    for data, want in VECTORS:
        assert K.fnv1a64(data) == want
        buf = C.create_string_buffer(data, len(data))
        assert lib.lbm_fnv1a64(buf, len(data)) == want


def test_spec_hash_separates_scenarios():
    base = dict(Q=19, dims=(16, 16, 16), nu=0.03, collision="trt", magic=3 / 16, omega_shear=None,
                periodic=(0, 0, 0), force=(0, 0, 0))
    h0 = K.spec_hash(**base)
    for key, val in [("nu", 0.0300000001), ("magic", 0.25), ("periodic", (1, 0, 0)), ("force", (1e-6, 0, 0)),
                     ("dims", (16, 16, 17)), ("collision", "bgk")]:
        assert K.spec_hash(**{**base, key: val}) != h0, key
    assert K.decomp_hash((16, 16, 16), (0, 0, 0)) != K.decomp_hash((16, 16, 16), (1, 0, 0))
