"""Host logic of the kernel-family table (include/cbk_ext.h,
cbk_set_kernel_family / cbk_get_kernel_family): argument codes and the
install / query / remove round trip. Nothing is launched."""

import cbk_amd

F = cbk_amd.FAMILY


def test_family_table_round_trip():
    key = ("d", "posv", 24, 24, 4, 64)
    assert cbk_amd.get_kernel_family(*key) == F["default"]
    assert cbk_amd.set_kernel_family(*key, F["team"]) == 0
    assert cbk_amd.get_kernel_family(*key) == F["team"]
    # another V, precision or nrhs is another entry
    assert cbk_amd.get_kernel_family("d", "posv", 24, 24, 4, 32) == F["default"]
    assert cbk_amd.get_kernel_family("s", "posv", 24, 24, 4, 64) == F["default"]
    assert cbk_amd.get_kernel_family("d", "posv", 24, 24, 2, 64) == F["default"]
    assert cbk_amd.set_kernel_family(*key, F["default"]) == 0
    assert cbk_amd.get_kernel_family(*key) == F["default"]
    # trsm: the forward and the backward sweep are separate entries
    assert cbk_amd.set_kernel_family("s", "trsm", 24, 24, 0, 64, F["quad"], uplo="L", trans="N") == 0
    assert cbk_amd.get_kernel_family("s", "trsm", 24, 24, 0, 64, uplo="U", trans="T") == F["quad"]
    assert cbk_amd.get_kernel_family("s", "trsm", 24, 24, 0, 64, uplo="U", trans="N") == F["default"]
    assert cbk_amd.set_kernel_family("s", "trsm", 24, 24, 0, 64, F["default"], uplo="L", trans="N") == 0


def test_family_argument_codes():
    assert cbk_amd.set_kernel_family("x", "posv", 8, 8, 1, 64, 1) == -1
    assert cbk_amd.set_kernel_family("d", "gesv", 8, 8, 1, 64, 1) == -2
    assert cbk_amd.set_kernel_family("d", "trsm", 8, 8, 0, 64, 1, uplo="X") == -3
    assert cbk_amd.set_kernel_family("d", "trsm", 0, 8, 0, 64, 1) == -5
    assert cbk_amd.set_kernel_family("d", "posv", 8, 65, 1, 64, 1) == -6
    assert cbk_amd.set_kernel_family("d", "gels", 8, 9, 1, 64, 1) == -6
    assert cbk_amd.set_kernel_family("d", "gels", 16, 8, 0, 64, 1) == -7
    assert cbk_amd.set_kernel_family("d", "posv", 8, 8, 1, 12, 1) == -8
    assert cbk_amd.set_kernel_family("d", "posv", 8, 8, 1, 64, 8) == -10
    assert cbk_amd.get_kernel_family("d", "geqrf", 4, 8, 0, 64) == -6
    assert cbk_amd.autotune_compact("d", "posv", 8, 8, 1, 12)[0] == -8
    assert cbk_amd.autotune_compact("d", "posv", 8, 8, 1, 64, nm=-1)[0] == -9
