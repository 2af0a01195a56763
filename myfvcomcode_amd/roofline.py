"""Algorithmic (compulsory) HBM bytes per kernel launch.

Bytes = distinct array entries a launch must read + entries it writes, with
static geometry and topology counted, intermediate arrays counted only when
they cross a kernel boundary.  One function per kernel family, in terms of
the mesh sizes, so that the benchmark and the documentation use the same
model.  FP64 reals are 8 bytes, indices 4 bytes.
"""

R8 = 8
I4 = 4


def sizes(model):
    h = model.hdr
    return dict(M=h.m, N=h.n, NE=h.ne, NCV=h.ncv, KB=h.kb, KBM1=h.kbm1,
                RSN=h.nring_sn, RVE=h.nring_ve)


def ext_elevation(s, real=R8):
    """k_ext_elevation: EXTEL_EDGE gather + elevation update, one launch."""
    M, N, NCV = s["M"], s["N"], s["NCV"]
    reads = (N * 3 * real                 # D1, UA, VA of the stage input
             + NCV * (I4 + 2 * real)      # NTRG, DLTXE, DLTYE
             + (M + 2) * I4 + 2 * NCV * I4  # node->sub-edge CSR
             + M * (3 * real + 2 * I4))   # ELRK, ART1, H; node OBC / river flags
    writes = M * 2 * real                 # ELF, D
    return reads + writes


def ext_momentum(s, real=R8):
    """k_ext_momentum: ADVAVE_EDGE_GCN gather + EXTUV_EDGE, one launch."""
    M, N, NE = s["M"], s["N"], s["NE"]
    elem_reals = (3        # UA, VA, D1 stage input
                  + 3      # UARK, VARK, ELRK1
                  + 6      # ADX2D, ADY2D, DRX2D, DRY2D, WUBOT, WVBOT
                  + 11     # ART, H1, COR, CC_HVC, CC_SPONGE, EPOR, GRAV_E, XC, YC, WSX, WSY
                  + 8)     # A1U(4), A2U(4)
    elem_ints = 3 + 3 + 3 + 2 + 1 + 1   # NV, NBE, edge list, ISBCE/IUCP, FO/river offsets
    node_reals = 3         # ELF (new), EL, D (stage input)
    edge = NE * (4 * I4 + I4 + 4 * real)  # IEC, IENODE, ISBC, XIJC, YIJC, DLTXC, DLTYC
    writes = N * 4 * real  # UAF, VAF, EL1, D1
    return N * (elem_reals * real + elem_ints * I4) + M * node_reals * real + edge + writes


def external_stage(s, real=R8):
    return ext_elevation(s, real) + ext_momentum(s, real)


FAMILIES = {
    "ext_elevation": ext_elevation,
    "ext_momentum": ext_momentum,
}


def bytes_per_launch(name, model, real=R8):
    return FAMILIES[name](sizes(model), real)
