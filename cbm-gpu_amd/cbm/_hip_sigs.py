"""ctypes declarations of include/cbm_hip.h."""
import ctypes

c_i32, c_i64, c_vp = ctypes.c_int32, ctypes.c_int64, ctypes.c_void_p

# enum cbm_mode / cbm_dtype / candidate paths
MODE_AX, MODE_ADX, MODE_DADX = 0, 1, 2
F32, BF16 = 0, 1
CAND_AUTO, CAND_TILES, CAND_LISTS = 0, 1, 2


class CbmDesc(ctypes.Structure):
    """struct cbm_desc (include/cbm_hip.h)."""
    _fields_ = [
        ("n_rows", c_i32),
        ("n_levels", c_i32),
        ("nnz", c_i64),
        ("n_edges", c_i64),
        ("mode", c_i32),
        ("lvl_ptr_host", c_vp),
        ("lvl_nlong_host", c_vp),
        ("fck_ptr_host", c_vp),
        ("fred_ptr_host", c_vp),
        ("rev_ptr_host", c_vp),
        ("lvl_node", c_vp),
        ("lvl_parent", c_vp),
        ("lvl_ppos", c_vp),
        ("fwd_ptr", c_vp),
        ("fwd_col", c_vp),
        ("rev_node", c_vp),
        ("ch_ptr", c_vp),
        ("ch_node", c_vp),
        ("ch_mid", c_vp),
        ("fck_pos", c_vp),
        ("fck_meta", c_vp),
        ("fck_slot", c_vp),
        ("fred_pos", c_vp),
        ("fred_slot", c_vp),
        ("fred_cnt", c_vp),
        ("fwd_slots", c_i32),
        ("bwd_node", c_vp),
        ("bwd_ptr", c_vp),
        ("bwd_row", c_vp),
        ("bwd_mid", c_vp),
        ("bck_pos", c_vp),
        ("bck_meta", c_vp),
        ("bck_slot", c_vp),
        ("bred_pos", c_vp),
        ("bred_slot", c_vp),
        ("bred_cnt", c_vp),
        ("bwd_nlong", c_i32),
        ("n_bck", c_i32),
        ("n_bred", c_i32),
        ("bwd_slots", c_i32),
        ("d", c_vp),
        ("fck_gslot", c_vp),
        ("fred_gslot", c_vp),
        ("fwd_gslots", c_i32),
        ("rev_nlong_host", c_vp),
        ("rck_ptr_host", c_vp),
        ("rred_ptr_host", c_vp),
        ("rck_pos", c_vp),
        ("rck_meta", c_vp),
        ("rck_slot", c_vp),
        ("rred_pos", c_vp),
        ("rred_slot", c_vp),
        ("rred_cnt", c_vp),
        ("rev_slots", c_i32),
        ("fpc_range_host", c_vp),
        ("fpc_b", c_vp),
        ("fpc_e", c_vp),
        ("fpc_slot", c_vp),
        ("frs_slot", c_vp),
        ("frs_cnt", c_vp),
        ("fwd_pieces", c_i32),
        ("bpc_range_host", c_vp),
        ("bpc_b", c_vp),
        ("bpc_e", c_vp),
        ("bpc_mid", c_vp),
        ("bpc_slot", c_vp),
        ("brs_slot", c_vp),
        ("brs_cnt", c_vp),
        ("bwd_pieces", c_i32),
        ("bpc_row", c_vp),
        ("fpc4_range_host", c_vp),
        ("fpc4_b", c_vp),
        ("fpc4_e", c_vp),
        ("fpc4_slot", c_vp),
        ("frs4_slot", c_vp),
        ("frs4_cnt", c_vp),
        ("fwd_pieces4", c_i32),
        ("bpc4_range_host", c_vp),
        ("bpc4_b", c_vp),
        ("bpc4_e", c_vp),
        ("bpc4_mid", c_vp),
        ("bpc4_slot", c_vp),
        ("brs4_slot", c_vp),
        ("brs4_cnt", c_vp),
        ("bwd_pieces4", c_i32),
        ("bpc4_row", c_vp),
        ("tf_level", c_i32),
        ("tf_frags", c_i32),
        ("tf_frag", c_vp),
        ("tf_seg", c_vp),
        ("tf_pos", c_vp),
        ("tr_level", c_i32),
        ("tr_frags", c_i32),
        ("tr_frag", c_vp),
        ("tr_seg", c_vp),
        ("tr_pos", c_vp),
        ("n_units", c_i32),
        ("unit_levels", c_i32),
        ("un_ptr", c_vp),
        ("un_row", c_vp),
        ("un_lpar", c_vp),
        ("un_pc", c_vp),
        ("un_pcb", c_vp),
        ("un_pce", c_vp),
        ("un_seg_ptr", c_vp),
        ("un_seg", c_vp),
        ("un_of", c_vp),
        ("un_waited", c_vp),
    ]

# Device arrays of the plan that the descriptor points at.
DEVICE_ARRAYS = ("lvl_node", "lvl_parent", "lvl_ppos", "fwd_ptr", "fwd_col", "rev_node", "ch_ptr", "ch_node", "ch_mid", "fck_pos", "fck_meta",
                 "fck_slot", "fred_pos", "fred_slot", "fred_cnt", "bwd_node", "bwd_ptr", "bwd_row", "bwd_mid", "bck_pos",
                 "bck_meta", "bck_slot", "bred_pos", "bred_slot", "bred_cnt", "fck_gslot", "fred_gslot", "rck_pos", "rck_meta", "rck_slot", "rred_pos",
                 "rred_slot", "rred_cnt", "fpc_b", "fpc_e", "fpc_slot", "frs_slot", "frs_cnt", "bpc_b", "bpc_e", "bpc_mid",
                 "bpc_slot", "brs_slot", "brs_cnt", "bpc_row", "fpc4_b", "fpc4_e", "fpc4_slot", "frs4_slot", "frs4_cnt",
                 "bpc4_b", "bpc4_e", "bpc4_mid", "bpc4_slot", "brs4_slot", "brs4_cnt", "bpc4_row", "tf_frag", "tf_seg",
                 "tf_pos", "tr_frag", "tr_seg", "tr_pos", "un_ptr", "un_row", "un_lpar", "un_pc", "un_pcb", "un_pce",
                 "un_seg_ptr", "un_seg", "un_of", "un_waited")
HOST_ARRAYS = ("lvl_ptr", "lvl_nlong", "fck_ptr", "fred_ptr", "rev_ptr", "rev_nlong", "rck_ptr", "rred_ptr", "fpc_range",
               "bpc_range", "fpc4_range", "bpc4_range")


def declare(lib, sig):
    P_vp = ctypes.POINTER(c_vp)
    sig(lib, "cbm_last_error", ctypes.c_char_p)
    sig(lib, "cbm_desc_size", c_i32)
    sig(lib, "cbm_create", c_i32, ctypes.POINTER(CbmDesc), c_vp, P_vp)
    sig(lib, "cbm_destroy", c_i32, c_vp)
    sig(lib, "cbm_reserve", c_i32, c_vp, c_i32)
    sig(lib, "cbm_matmul", c_i32, c_vp, c_i32, c_vp, c_i64, c_vp, c_i64, c_i32, c_vp)
    sig(lib, "cbm_matmul_noalloc", c_i32, c_vp, c_i32, c_vp, c_i64, c_vp, c_i64, c_i32, c_vp)
    sig(lib, "cbm_matmul_t", c_i32, c_vp, c_i32, c_vp, c_i64, c_vp, c_i64, c_vp, c_i64, c_i32, c_vp)
    sig(lib, "cbm_adjoint_reverse", c_i32, c_vp, c_i32, c_vp, c_i64, c_vp, c_i64, c_i32, c_vp)
    sig(lib, "cbm_adjoint_delta_t", c_i32, c_vp, c_i32, c_vp, c_i64, c_vp, c_i64, c_vp, c_i64, c_i32, c_vp)
    sig(lib, "cbm_delta_spmm", c_i32, c_vp, c_i32, c_vp, c_i64, c_vp, c_i64, c_i32, c_vp)
    sig(lib, "cbm_propagate", c_i32, c_vp, c_vp, c_i64, c_i32, c_vp)
    sig(lib, "cbm_last_launches", c_i32, c_vp)
    sig(lib, "cbm_set_slab_bytes", c_i32, c_vp, c_i64)
    sig(lib, "cbm_xcd_ranges_active", c_i32, c_vp, c_i32, c_i32)
    sig(lib, "cbm_set_xcd_phases", c_i32, c_vp, c_i32)
    sig(lib, "cbm_set_tree_tails", c_i32, c_vp, c_i32)
    sig(lib, "cbm_set_symmetric", c_i32, c_vp, c_i32)
    sig(lib, "cbm_set_fused", c_i32, c_vp, c_i32)
    sig(lib, "cbm_attach_uncompressed", c_i32, c_vp, c_vp, c_i32)
    sig(lib, "cbm_uncompressed_active", c_i32, c_vp, c_i32, c_i32)
    sig(lib, "cbm_attach_order", c_i32, c_vp, c_i32, c_vp, c_i32, c_vp)
    sig(lib, "cbm_set_order", c_i32, c_vp, c_i32)
    sig(lib, "cbm_order_active", c_i32, c_vp, c_i32, c_i32)
    sig(lib, "cbm_fused_active", c_i32, c_vp, c_i32, c_i32)
    sig(lib, "cbm_fused_errors", c_i64, c_vp)
    sig(lib, "cbm_set_kahan_depth", c_i32, c_vp, c_i32)
    sig(lib, "cbm_set_forward_split", c_i32, c_vp, c_i32)
    sig(lib, "cbm_set_wave_rows", c_i32, c_vp, c_i32)
    sig(lib, "cbm_set_mixed_grid", c_i32, c_vp, c_i32)
    sig(lib, "cbm_set_chunk_finish", c_i32, c_vp, c_i32)
    sig(lib, "cbm_set_xcd_slabs", c_i32, c_vp, c_i32)
    sig(lib, "cbm_set_xcd_ranges", c_i32, c_vp, c_i32)
    sig(lib, "csr_last_error", ctypes.c_char_p)
    sig(lib, "csr_create", c_i32, c_i32, c_i32, c_i64, c_vp, c_vp, c_vp, P_vp)
    sig(lib, "csr_destroy", c_i32, c_vp)
    sig(lib, "csr_set_algorithm", c_i32, c_vp, c_i32)
    sig(lib, "csr_matmul", c_i32, c_vp, c_i32, c_vp, c_i64, c_vp, c_i64, c_i32, c_vp)
    sig(lib, "cbm_op_last_error", ctypes.c_char_p)
    sig(lib, "cbm_op_spmm", c_i32, c_i32, c_vp, c_vp, c_vp, c_vp, c_vp, c_i64, c_vp, c_i64, c_i32, ctypes.c_float,
        c_i32, c_vp)
    sig(lib, "cbm_op_update", c_i32, c_i32, c_vp, c_vp, c_vp, c_vp, c_vp, c_i64, c_i32, c_vp)
    sig(lib, "cbm_op_spmm_t_workspace_bytes", c_i64, c_i32, c_i32, c_i64)
    sig(lib, "cbm_op_spmm_t", c_i32, c_i32, c_i32, c_vp, c_vp, c_vp, c_vp, c_i64, c_vp, c_i64, c_vp, c_i64, c_i32, c_vp,
        c_i64, c_vp)
    sig(lib, "cbm_op_axpy", c_i32, c_i32, c_i32, ctypes.c_float, c_vp, c_i64, c_vp, c_i64, c_vp)
    sig(lib, "cbm_candidates_last_error", ctypes.c_char_p)
    sig(lib, "cbm_cluster_last_error", ctypes.c_char_p)
    sig(lib, "cbm_cluster_column_counts", c_i32, c_i32, c_i32, c_vp, c_vp, c_vp, c_i32, c_vp, c_vp)
    sig(lib, "cbm_cluster_fingerprints", c_i32, c_i32, c_vp, c_vp, c_vp, c_i32, c_vp, c_i32, c_vp, c_vp)
    sig(lib, "cbm_candidates_workspace_bytes", c_i64, c_i32, c_i32, c_i32)
    sig(lib, "cbm_candidates_count", c_i32, c_i32, c_vp, c_vp, c_vp, c_vp, c_i32, c_i32, c_i32, c_vp, c_vp, c_i64,
        c_vp)
    sig(lib, "cbm_candidates_emit", c_i32, c_i32, c_vp, c_vp, c_vp, c_vp, c_i32, c_i32, c_i32, c_vp, c_vp, c_vp,
        c_vp, c_i64, c_vp, c_i64, c_vp)
    sig(lib, "cbm_partners_workspace_bytes", c_i64, c_i32, c_i32)
    sig(lib, "cbm_partners", c_i32, c_i32, c_vp, c_vp, c_i32, c_vp, c_vp, c_i32, c_i32, c_i32, c_vp, c_vp, c_vp, c_i64,
        c_vp)


# Every symbol include/cbm_hip.h declares (checked by the CPU test suite).
EXPORTED = ("cbm_last_error", "cbm_desc_size", "cbm_create", "cbm_destroy", "cbm_reserve", "cbm_matmul", "cbm_matmul_noalloc", "cbm_matmul_t", "cbm_adjoint_reverse",
            "cbm_adjoint_delta_t", "cbm_delta_spmm",
            "cbm_propagate", "cbm_last_launches", "cbm_set_slab_bytes", "cbm_xcd_ranges_active", "cbm_set_xcd_phases", "cbm_set_tree_tails", "cbm_set_symmetric", "cbm_set_fused", "cbm_attach_uncompressed", "cbm_uncompressed_active", "cbm_attach_order", "cbm_set_order", "cbm_order_active", "cbm_fused_active", "cbm_fused_errors", "cbm_set_kahan_depth",
            "cbm_set_forward_split", "cbm_set_wave_rows", "cbm_set_mixed_grid", "cbm_set_chunk_finish", "cbm_set_xcd_slabs", "cbm_set_xcd_ranges", "csr_create", "csr_destroy", "csr_matmul", "csr_set_algorithm",
            "csr_last_error", "cbm_op_spmm", "cbm_op_spmm_t_workspace_bytes", "cbm_op_spmm_t", "cbm_op_update", "cbm_op_axpy", "cbm_op_last_error",
            "cbm_candidates_workspace_bytes", "cbm_candidates_count", "cbm_candidates_emit",
            "cbm_partners_workspace_bytes", "cbm_partners",
            "cbm_candidates_last_error", "cbm_cluster_column_counts", "cbm_cluster_fingerprints",
            "cbm_cluster_last_error")
