"""Drop-in for the reference's ``eob.pytdinterp`` Cython module (src/tdinterp.pyx).

Array arguments are GPU arrays; complex arrays are complex128.  Calls are
synchronous like the reference's wraps.
"""
from __future__ import annotations

from ._lib import call
from .utils import current_stream, pointer_adjust, sync

__all__ = ["interpolate_TD_wrap", "all_in_one_likelihood", "TDInterp_wrap2", "TDInterp_wrap"]


@pointer_adjust
def interpolate_TD_wrap(t_arr, propArrays, B, upper_diag, diag, lower_diag, lengths, numBinAll, nsubs):
    """src/tdinterp.pyx:20-37 -> interpolate_TD (src/interpolate.cu:269-308).

    Outputs alias as in the reference: upper_diag <- c1, diag <- c2, lower_diag <- c3.
    """
    call("eobhip_interpolate_TD", t_arr, propArrays, B, upper_diag, diag, lower_diag, lengths, int(numBinAll), int(nsubs),
         current_stream())
    sync()


@pointer_adjust
def all_in_one_likelihood(temp_sum, hp_fd, hc_fd, data, psd, Fplus, Fcross, time_shift, df, num_bin_all, nchannels, data_length):
    """src/tdinterp.pyx:70-88 -> all_in_one_likelihood_wrap (src/interpolate.cu:616-628)."""
    call("eobhip_all_in_one_likelihood", temp_sum, hp_fd, hc_fd, data, psd, Fplus, Fcross, time_shift, float(df), int(num_bin_all),
         int(nchannels), int(data_length), current_stream())
    sync()


@pointer_adjust
def TDInterp_wrap2(templateChannels, dataTime, ts, propArrays, c1, c2, c3, old_length, old_lengths, data_length, numBinAll, numModes,
                   ls, ms, inds, lengths, max_length, legacy_stride=True):
    """src/tdinterp.pyx:60-75 -> TDInterp2 (src/interpolate.cu:441-549): the mode sum
    sum_j (re_j + i im_j) e^{-i m_j phi} at dataTime[i], i < lengths[b], into
    templateChannels [numBinAll][data_length] (complex128).

    legacy_stride=True indexes binary b's series with stride old_lengths[b], as the
    reference does (src/interpolate.cu:521-527), which is right only when every
    binary has old_length knots; False uses the padded stride old_length (the
    layout [numBinAll][2 numModes + 1][old_length] the arrays have).  ls, ms, inds,
    lengths, old_lengths are int32 device arrays."""
    call("eobhip_TDInterp2", templateChannels, dataTime, ts, propArrays, c1, c2, c3, int(old_length), old_lengths, int(data_length),
         int(numBinAll), int(numModes), ls, ms, inds, lengths, int(max_length), 1 if legacy_stride else 0, current_stream())
    sync()


def TDInterp_wrap(*args, **kwargs):
    """src/tdinterp.pyx:39-53 (pointer-array TD, unused by the reference)."""
    raise NotImplementedError("TDInterp_wrap is dead code in the reference and is not provided")
