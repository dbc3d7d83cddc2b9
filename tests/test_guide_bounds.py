"""Host-side bounds check of the guide kernels' indexing (k_guide.hip).

Every fragment load, bias load, activation / delta store and LDS staging
access of the inference, training and weight-gradient kernels is replayed
over all lanes, tiles and layers at hidden widths 64 and 128, and must fall
inside the byte size the kernel gives its buffer resource (make_rsrc) or the
host allocates (guide_capi.cpp: init_guide, ensure_batch, guide_prepare).
Raw buffer accesses past a resource's size do not fault on gfx950: loads read
zero and stores are dropped, so an extent error there would corrupt the
network silently rather than stop the kernel.  The int32 resource sizes and
store offsets must stay below 2^31 at the largest batch the host accepts.

The layer geometry restates GuideDims (tracer_amd/csrc/tr_guide.h); the
access expressions are those of the kernels, cited by function."""
import pytest

IN, OUT, LAYERS = 96, 48, 5


class Dims:
    """GuideDims (tr_guide.h)."""

    def __init__(self, H):
        self.H = H

    def inp(self, l):
        return IN if l == 0 else self.H

    def out(self, l):
        return OUT if l == LAYERS - 1 else self.H

    def out_k(self, l):
        return (self.out(l) + 31) // 32 * 32

    def fwd_offset(self, l):
        return sum(self.inp(i) * self.out(i) for i in range(l))

    def bwd_offset(self, l):
        return sum(self.inp(i) * self.out_k(i) for i in range(l))

    def fwd_size(self):
        return self.fwd_offset(LAYERS)

    def bwd_size(self):
        return self.bwd_offset(LAYERS)

    def act_rows(self):
        return IN + 4 * self.H

    def delta_rows(self):
        return 4 * self.H + OUT

    def act_row(self, l):
        return 0 if l == 0 else IN + (l - 1) * self.H

    def delta_row(self, l):
        return l * self.H


def slot_row(j):
    return j if j < 4 else 12 + j


def mlp_layer_extent(KS, MT, base):
    """Largest byte + 1 that mlp_layer's fragment loads touch: fragment
    (mt, s) at base + (mt * KS + s) * 1024, 64 lanes x 16 bytes."""
    hi = 0
    for mt in range(MT):
        for s in range(KS):
            for lane in (0, 63):
                hi = max(hi, base + (mt * KS + s) * 1024 + lane * 16 + 16)
    return hi


@pytest.mark.parametrize("H", [64, 128])
def test_forward_and_backward_fragments_in_bounds(H):
    d = Dims(H)
    fwd_bytes, bwd_bytes = d.fwd_size() * 2, d.bwd_size() * 2
    # forward_hidden / forward_all: layer 0 (3 K steps), layers 1..3 (H/32), head (H/32 steps, 3 tiles)
    ext = [mlp_layer_extent(3, H // 16, d.fwd_offset(0) * 2)]
    ext += [mlp_layer_extent(H // 32, H // 16, d.fwd_offset(l) * 2) for l in (1, 2, 3)]
    ext.append(mlp_layer_extent(H // 32, 3, d.fwd_offset(4) * 2))
    assert max(ext) == fwd_bytes   # the last fragment ends exactly at the buffer's end
    # k_guide_train backward: head (2 K steps over 48 outputs padded to 64), then layers 3..1
    bext = [mlp_layer_extent(2, H // 16, d.bwd_offset(4) * 2)]
    bext += [mlp_layer_extent(H // 32, H // 16, d.bwd_offset(l) * 2) for l in (3, 2, 1)]
    assert max(bext) <= bwd_bytes
    # every layer's fragments are disjoint and in order
    for l in range(LAYERS):
        assert d.fwd_offset(l) * 2 % 1024 == 0 and d.bwd_offset(l) * 2 % 1024 == 0


@pytest.mark.parametrize("H", [64, 128])
def test_bias_loads_in_bounds(H):
    bias_bytes = (4 * H + OUT) * 4
    hi = 0
    for l in range(4):             # pack_relu: bias4(l H + 32 s + {0, 16}, g)
        for s in range(H // 32):
            for off in (0, 16):
                for g in range(4):
                    hi = max(hi, (l * H + 32 * s + off) * 4 + g * 16 + 16)
    assert hi == 4 * H * 4
    for mt in range(3):            # forward_all: output bias4(4 H + 16 mt, g)
        for g in range(4):
            hi = max(hi, (4 * H + mt * 16) * 4 + g * 16 + 16)
    assert hi == bias_bytes


@pytest.mark.parametrize("H, limit", [(64, 64 * 1024), (128, 160 * 1024)])
def test_inference_lds_staging_fits(H, limit):
    """k_guide_infer<H, NT, true, BLOCK> stages fwd_size / 8 fragments of 16
    bytes in the dynamic LDS that launch_guide_infer passes (fwd_size * 2
    bytes): width 64 within 64 KB (three workgroups on a CU's 160 KB), width
    128 within the 160 KB a gfx950 workgroup may hold (the launcher checks the
    device's limit and otherwise reads the fragments from global memory)."""
    d = Dims(H)
    staged = d.fwd_size() // 8 * 16
    assert d.fwd_size() % 8 == 0 and staged == d.fwd_size() * 2
    assert staged <= limit
    if H == 64:
        assert 3 * staged <= 160 * 1024


@pytest.mark.parametrize("H", [64, 128])
def test_training_activation_and_delta_stores_in_bounds(H):
    """k_guide_train's store_u16(acts | deltas, v, voff, soff) at byte
    voff + soff, voff = (4 g n_pad + idx) * 2, over every slot of every
    stored layer."""
    d = Dims(H)
    n_pad = 4096
    acts_rows, delta_rows = set(), set()
    for g in range(4):
        for j in range(8):
            for s in range(3):
                acts_rows.add(4 * g + 32 * s + slot_row(j))
            for l in range(4):
                for s in range(H // 32):
                    acts_rows.add(4 * g + d.act_row(l + 1) + 32 * s + slot_row(j))
                    delta_rows.add(4 * g + d.delta_row(l) + 32 * s + slot_row(j))
            delta_rows.add(4 * g + d.delta_row(4) + slot_row(j))
            if j < 4:
                delta_rows.add(4 * g + d.delta_row(4) + 32 + j)
    # the stored rows are exactly the activation / delta rows, each once per (g, slot)
    assert acts_rows == set(range(d.act_rows()))
    assert delta_rows == set(range(d.delta_rows()))
    for rows, total in ((acts_rows, d.act_rows()), (delta_rows, d.delta_rows())):
        hi = max((r * n_pad + n_pad - 1) * 2 + 2 for r in rows)
        assert hi == total * n_pad * 2


@pytest.mark.parametrize("H", [64, 128])
def test_weight_gradient_reads_in_bounds(H):
    """k_guide_dw reads rows act_row(l) + 16 mt + (0..15) and
    delta_row(l) + 16 nt + (0..15), samples k + 4 g .. k + 31 for k a
    multiple of 32 below n_pad (a multiple of 32)."""
    d = Dims(H)
    for l in range(LAYERS):
        assert d.inp(l) % 16 == 0 and d.out(l) % 16 == 0
        assert d.act_row(l) + d.inp(l) <= d.act_rows()
        assert d.delta_row(l) + d.out(l) <= d.delta_rows()
    n_pad = 96
    for k in range(0, n_pad, 32):
        assert k + 16 + 4 * 3 + 3 < n_pad


def test_int32_sizes_at_the_largest_accepted_batch():
    """check_config rejects act_rows * n_pad * 2 >= 2^31; below that bound every
    resource size and store offset of the training kernel fits an int."""
    for H in (64, 128):
        d = Dims(H)
        n_max = ((1 << 31) - 1) // (d.act_rows() * 2) // 32 * 32
        assert d.act_rows() * n_max * 2 < (1 << 31)
        assert d.delta_rows() * n_max * 2 < (1 << 31)   # delta_rows <= act_rows
        assert d.delta_rows() <= d.act_rows()
