"""The wire format of device data: a ciphertext, a plaintext and a key
written from the device and read back are bit-identical with their forms and
long double scale; corrupted payloads, foreign prime chains, truncated
records and small buffers are refused."""
import numpy as np
import pytest

import fhe_amd
from support import uniform_limbs

pytestmark = pytest.mark.gpu


def test_wire_round_trip_and_refusals(gpu_lib):
    with fhe_amd.Context(14, 2) as ctx:
        rng = np.random.Generator(np.random.PCG64(4))
        n, mid = ctx.n, 2
        limbs = ctx.limbs_at(mid)
        scale = np.longdouble(2.0) ** 80 / np.longdouble(3.0)
        for kind, polys, level, flags in [
                (fhe_amd.WIRE["CIPHERTEXT"], 2, 5, 0),
                (fhe_amd.WIRE["CIPHERTEXT"], 3, 4, fhe_amd.FLAG_NTT_FORM),
                (fhe_amd.WIRE["PLAINTEXT"], 1, 5, fhe_amd.FLAG_RESCALED),
                (fhe_amd.WIRE["KEY"], 4, -1, fhe_amd.FLAG_NTT_FORM)]:
            x = np.stack([uniform_limbs(rng, ctx.primes, mid, limbs, n)
                          for _ in range(polys)])
            x[0, 0, :3] = [-5, 0, -(1 << 40)]  # signed representatives too
            d = ctx.to_device(x)
            blob = ctx.wire_write(kind, d, mid, level, flags, scale)
            h = ctx.wire_header(blob)
            assert (h["kind"], h["polys"], h["moduli_id"], h["limbs"],
                    h["level"], h["flags"]) == (kind, polys, mid, limbs,
                                                level, flags)
            assert h["scale"] == scale
            out = ctx.array(x.shape)
            ctx.wire_read(blob, out)
            assert np.array_equal(out.download(), x)
            # a flipped payload bit, a truncated record, a foreign chain
            bad = bytearray(blob)
            bad[-1] ^= 1
            with pytest.raises(fhe_amd.FheError):
                ctx.wire_read(bytes(bad), out)
            with pytest.raises(fhe_amd.FheError):
                ctx.wire_read(blob[:-8], out)
            with pytest.raises(fhe_amd.FheError):
                ctx.wire_header(b"\0" * 80)
            small = ctx.array((1, 1, n))
            with pytest.raises(fhe_amd.FheError):
                ctx.wire_read(blob, small)
            for arr in (d, out, small):
                arr.free()
        # forged headers (the checksum covers the payload only): a limb
        # range past the chain whose int sum would wrap, a polynomial count
        # whose byte size would wrap, a level off the chain, unknown flags
        import struct
        for off, value in [(32, 2 ** 31 - 1), (36, 2 ** 30), (28, 2 ** 15),
                           (28, 2 ** 31 - 1), (40, 1000), (40, -2),
                           (44, 8)]:
            forged = bytearray(blob)
            forged[off:off + 4] = struct.pack("<i", value)
            with pytest.raises(fhe_amd.FheError):
                ctx.wire_header(bytes(forged))
        with fhe_amd.Context(13, 1) as other:
            with pytest.raises(fhe_amd.FheError):
                other.wire_header(blob)
