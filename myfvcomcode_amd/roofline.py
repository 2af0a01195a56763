"""Algorithmic (compulsory) HBM bytes per kernel launch.

Bytes = distinct array entries a launch must read + entries it writes.
Static geometry and topology are counted; intermediate arrays are counted
only where they cross a kernel boundary in this design (the per-sub-edge and
per-edge flux buffers, the node-layer records).  One function per kernel
family (the names the library's timing events use), in terms of the mesh
sizes, so that the benchmark and the documentation use the same model.
FP64 reals are 8 bytes, indices 4 bytes; `real` is the size of the state
reals (8 for the FP64 path).

Per-stage conditional traffic (EGF at RK stage 3, UARD at stage 3, the
OBC and river branches) is left out: it is small and not every launch has it.
"""

R8 = 8
I4 = 4


def sizes(model):
    h = model.hdr
    return dict(M=h.m, N=h.n, NE=h.ne, NCV=h.ncv, KB=h.kb, KBM1=h.kbm1,
                RSN=h.nring_sn, RVE=h.nring_ve)


# ----------------------------------------------------------------- 2-D mode
def ext_elevation(s, real=R8):
    """EXTEL_EDGE gather + elevation update (one RK stage)."""
    M, N, NCV = s["M"], s["N"], s["NCV"]
    reads = (M * (3 * I4 + 3 * real)          # cv_cnt, node_obc, node_riv; ELRK, ART1, H
             + 2 * NCV * I4                   # slot-major sub-edge lists (term position, side)
             + NCV * real)                    # the stage input's EXTEL term of every sub-edge
    writes = M * 2 * real                     # EL, D of the stage output
    return reads + writes


def advave_edge(s, real=R8):
    """ADVAVE_EDGE_GCN edge terms, one thread per edge."""
    M, N, NE = s["M"], s["N"], s["NE"]
    reads = (NE * (4 * I4 + 4 * real)         # IEC, IENODE; XIJC, YIJC, DLTXC, DLTYC
             + N * (3 * I4 + 8 * real + 4 * real + 2 * real)  # NBE; A1U, A2U; XC, YC, ART, CC_HVC; UA, VA
             + M * 2 * real)                  # D, EL
    writes = NE * (4 * real + real)           # EdgeFlux4 + ELIJ
    return reads + writes


def ext_momentum(s, real=R8):
    """Edge-term gather + EXTUV_EDGE + wall projection (one RK stage)."""
    M, N, NE = s["M"], s["N"], s["NE"]
    elem_reals = 20   # EPOR GRAV_E D1 UA VA ART COR WSX WSY ADX2D ADY2D DRX2D DRY2D WUBOT WVBOT H1 EL1 UARK VARK CC_SPONGE
    elem_ints = 9     # NV(3), edge list(3), IUCP, ISBCE, river offset
    reads = (N * (elem_reals * real + elem_ints * I4)
             + NE * (4 * real + real + I4 + 2 * real)  # EdgeFlux4, ELIJ, ISBC, DLTXC, DLTYC
             + N * 6 * real                            # DLTXE/DLTYE of the element's 3 sub-edges
             + M * real)                               # new EL
    writes = N * 4 * real + N * 3 * real               # UA, VA, EL1, D1; EXTEL terms of the next stage
    return reads + writes


# ----------------------------------------------------------------- 3-D mode
def _cvrec(s, layers, nfields, real):
    M, RSN = s["M"], s["RSN"]
    reads = RSN * (I4 + 2 * real) + M * (2 * I4 + 2 * real) + nfields * M * layers * real
    writes = M * layers * 8 * real
    return reads + writes


def cvrec_ts(s, real=R8):
    """Ring gradients/bounds + viscosity per node-layer (T or S, KBM1 layers)."""
    return _cvrec(s, s["KBM1"], 3, real)       # f, mean, VISCOFH


def cvrec_q(s, real=R8):
    """The same for Q2 / Q2L on interfaces 2..KBM1."""
    return _cvrec(s, s["KBM1"] - 1, 2, real)   # q, VISCOFH


def _cvflux(s, rec_layers, uv_layers, real):
    M, N, NCV = s["M"], s["N"], s["NCV"]
    reads = (NCV * (3 * I4 + 6 * real)          # NIEC(2), NTRG; DLTXE, DLTYE, DLTXNCVE(2), DLTYNCVE(2)
             + M * real + N * real              # NN_HVC, DT1
             + M * rec_layers * 8 * real        # node-layer records
             + N * uv_layers * 2 * real)        # U, V
    writes = NCV * rec_layers * real            # face flux
    return reads + writes


def cvflux_ts(s, real=R8):
    """ADV_T/ADV_S face flux per sub-edge and layer."""
    return _cvflux(s, s["KBM1"], s["KBM1"], real)


def cvflux_q(s, real=R8):
    """ADV_Q face flux per sub-edge and interface."""
    return _cvflux(s, s["KBM1"] - 1, s["KBM1"], real)


def _cvrecflux(s, layers, nfields, uv_layers, bottom_records, real):
    """Records formed in LDS and face fluxes in one launch: the node fields
    and ring geometry, the element geometry and velocities are read, the
    face fluxes (and the bottom layer's records for the tracer) written;
    no node-layer record crosses HBM."""
    M, N, NCV, RSN = s["M"], s["N"], s["NCV"], s["RSN"]
    reads = (RSN * (I4 + 2 * real) + M * (2 * I4 + 2 * real)   # ring lists and factors, NTSN, ART2, NN_HVC
             + nfields * M * layers * real                     # field, mean, VISCOFH
             + 2 * N * I4 + M * 13 // 10 * I4                  # block element lists and local indices; node lists (about 1.3 per node)
             + NCV * (3 * I4 + 6 * real) + N * real            # sub-edge geometry, DT1
             + N * uv_layers * 2 * real)                       # U, V
    writes = NCV * layers * real + (M * 8 * real if bottom_records else 0)
    return reads + writes


def cvrecflux_ts(s, real=R8):
    """ADV_T/ADV_S records and face fluxes in one launch."""
    return _cvrecflux(s, s["KBM1"], 3, s["KBM1"], True, real)


def cvrecflux_q(s, real=R8):
    """ADV_Q records and face fluxes in one launch."""
    return _cvrecflux(s, s["KBM1"] - 1, 2, s["KBM1"], False, real)


def adv_ts(s, real=R8):
    """Ordered sub-edge sums + vertical flux + update (T or S)."""
    M, NCV, KB, KBM1 = s["M"], s["NCV"], s["KB"], s["KBM1"]
    reads = (M * (3 * I4 + 4 * real) + 2 * NCV * I4   # counts, flags; ART1, DT, DTFA, NN_HVC; slot lists
             + NCV * KBM1 * real                       # face fluxes
             + M * KBM1 * real + M * KB * real         # T1, WTS
             + M * KBM1 * 2 * real                     # FCT ring bounds (RLOHI)
             + M * (real + 8 * real))                  # bottom VISCOFH and record
    writes = M * KBM1 * real + 3 * M * real            # TF1; PFPXB, PFPYB, AH_BOTTOM
    return reads + writes


def adv_q(s, real=R8):
    """Ordered sub-edge sums + vertical flux + update (Q2 or Q2L)."""
    M, NCV, KB, KBM1 = s["M"], s["NCV"], s["KB"], s["KBM1"]
    reads = (M * (I4 + 3 * real) + 2 * NCV * I4 + NCV * (KBM1 - 1) * real + 2 * M * KB * real
             + M * (KBM1 - 1) * 2 * real)              # FCT ring bounds (RLOHI)
    writes = M * KB * real
    return reads + writes


def _uv_blk(s, real, uv):
    """ADVECTION / ADV_UV by element block: the element and edge data, the
    velocities of each layer once, and the outputs; records and edge terms
    stay in LDS."""
    M, N, NE, KB, KBM1 = s["M"], s["N"], s["NE"], s["KB"], s["KBM1"]
    reads = (NE * (5 * I4 + 4 * real)                  # IEC, IENODE, ISBC; XIJC, YIJC, DLTXC, DLTYC
             + N * (3 * I4 + 12 * real)                # NBE; A1U, A2U, XC, YC, ART, CC_HVC
             + N * (4 * I4 + 2 * real)                 # edge list, IUCP/ISBCE, river offset; EPOR, ...
             + N * 7 * I4                              # block tables: records, edges (about two per element each), edge slots
             + M * real                                # DT
             + N * KBM1 * 2 * real)                    # U, V
    if uv:
        reads += (N * 5 * real + M * real               # DT1, D1, COR, GRAV_E, ART; EGF
                  + N * KBM1 * 2 * real + N * KB * real)  # DRHOX, DRHOY; W
        writes = N * KBM1 * 3 * real + N * real        # UF, VF, VISCOFM
    else:
        writes = 2 * N * real                          # ADX2D, ADY2D
    return reads + writes


def adv3d_blk(s, real=R8):
    """ADVECTION_EDGE_GCN by element block, summed into ADX2D/ADY2D."""
    return _uv_blk(s, real, False)


def advuv_blk(s, real=R8):
    """ADV_UV_EDGE_GCN by element block -> UF, VF, VISCOFM."""
    return _uv_blk(s, real, True)


def vertvl(s, real=R8):
    """VERTVL_EDGE by node tile: the elements' sub-edge geometry and U, V,
    the nodes' sub-edge positions, column integral, correction."""
    M, N, NCV, KB, KBM1 = s["M"], s["N"], s["NCV"], s["KB"], s["KBM1"]
    reads = (2 * NCV * I4 + M * (3 * I4 + 6 * real)   # sub-edge positions; counts, flags, ART1, D, DT, ELF, H
             + N * (I4 + 7 * real)                    # tile element lists; DT1, DLTXE/DLTYE of 3 sub-edges
             + N * KBM1 * 2 * real)                   # U, V
    writes = M * KB * real + M * real                 # WTS, DTFA
    return reads + writes


def viscofh(s, real=R8):
    """VISCOF_H: nodal strain rate from the element ring."""
    M, N, KBM1, RVE = s["M"], s["N"], s["KBM1"], s["RVE"]
    reads = RVE * (I4 + 4 * real) + M * (3 * I4 + real) + N * KBM1 * 2 * real
    writes = M * KBM1 * real
    return reads + writes


def vdif_q(s, real=R8):
    """VDIF_Q (MY2.5): column solves for Q2, Q2L and the mixing coefficients,
    then Q2 = Q2F, Q2L = Q2LF.  Reads Q2 Q2L RHO1 KQ KM KH Q2F Q2LF and the
    ring of U/V; writes Q2F Q2LF Q2 Q2L L KM KH KQ.  The in-kernel scratch
    (elimination coefficients, VHP parked in Q2F/Q2LF) is not counted."""
    M, N, KB, KBM1, RVE = s["M"], s["N"], s["KB"], s["KBM1"], s["RVE"]
    reads = M * KB * 8 * real + N * KBM1 * 2 * real + RVE * I4 + M * 4 * real + N * 4 * real
    writes = M * KB * 8 * real
    return reads + writes


def vdif_uv(s, real=R8):
    """VDIF_UV: implicit vertical diffusion of UF and VF."""
    N, KB, KBM1 = s["N"], s["KB"], s["KBM1"]
    return N * KB * 3 * real + N * 6 * real + N * KBM1 * 2 * real + N * 3 * real


def vdif_ts(s, real=R8):
    """VDIF_TS: implicit vertical diffusion of one tracer."""
    M, KB, KBM1 = s["M"], s["KB"], s["KBM1"]
    return M * KB * 2 * real + M * 8 * real + M * KBM1 * real


def n2e3d(s, real=R8):
    """N2E3D of one node-layer field (KB layers): element mean of its nodes."""
    M, N, KB = s["M"], s["N"], s["KB"]
    return M * KB * real + N * 3 * I4 + N * KB * real


FAMILIES = {f.__name__: f for f in (
    ext_elevation, advave_edge, ext_momentum, cvrec_ts, cvrec_q, cvflux_ts, cvflux_q, cvrecflux_ts, cvrecflux_q,
    adv_ts, adv_q,
    adv3d_blk, advuv_blk, vertvl, viscofh, vdif_q, vdif_uv, vdif_ts,
    n2e3d)}


def bytes_per_launch(name, model, real=R8):
    return FAMILIES[name](sizes(model), real)
