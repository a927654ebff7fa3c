"""Mode coupling of pseudo-C_l by a window -- drop-in for the reference's
``sht.mask_deconvolution.MaskDeconvolution`` (sht/mask_deconvolution.py:14-248).

The O(Nl^3) part, the mode-coupling matrix M_{l1 l2} (get_M), runs on the
device (``dsht_mode_coupling``, HIP kernel in libdsht.so); it needs no 3j
table.  The binning, the inversion of the small bin-by-bin matrix and the
products with it are the reference's host-side numpy steps.
"""
import numpy as np
import torch

from . import _lib
from .threej000 import Wigner3j


def mode_coupling_device(Nl, W, device=None, stream=None):
    """M[l1, l2] = (2 l2 + 1)/(4 pi) sum_l3 (2 l3 + 1) W[l3] (l1 l2 l3; 0 0 0)^2
    as an (Nl, Nl) float64 CUDA tensor; W (numpy or tensor) is taken as zero
    past its end."""
    if device is None:
        device = torch.device("cuda", torch.cuda.current_device())
    w = W if isinstance(W, torch.Tensor) else torch.from_numpy(np.ascontiguousarray(W, np.float64))
    w = w.to(device=device, dtype=torch.float64).contiguous().reshape(-1)
    lib = _lib.load()
    with torch.cuda.device(device):
        M = torch.empty((int(Nl), int(Nl)), dtype=torch.float64, device=device)
        s = stream if stream is not None else torch.cuda.current_stream(device)
        _lib.check(lib.dsht_mode_coupling(int(Nl), w.data_ptr() if w.numel() else None,
                                          w.numel(), M.data_ptr(), s.cuda_stream),
                   "dsht_mode_coupling")
    return M


class MaskDeconvolution:
    def __init__(self, Nl, W_l, verbose=True, precompute_3j_and_Mll=True, device=None):
        """sht/mask_deconvolution.py:15-52.  Nl multipoles (lmax = Nl - 1);
        W_l the window's power spectrum, zero-padded to 2 Nl - 1 values.
        The 3j table (``w3j000``, Wigner3j(2 Nl - 1)) is built on first
        access: the mode-coupling kernel does not need it."""
        self.Nl = int(Nl)
        self.lmax = self.Nl - 1
        W_l = np.asarray(W_l, dtype=np.float64)
        self.W_l = np.pad(W_l, (0, max(0, 2 * self.Nl - 1 - W_l.size)), "constant",
                          constant_values=0)
        self.device = device    # None: the current device when the GPU is used
        self._w3j000 = None
        self.verbose = verbose
        if precompute_3j_and_Mll:
            if verbose:
                print("Computing the mode-coupling matrix...")
            self.Mll = self.get_M()
        else:
            self.Mll = None

    @property
    def w3j000(self):
        """Wigner3j(2 Nl - 1) (sht/mask_deconvolution.py:42), built on first use."""
        if self._w3j000 is None:
            self._w3j000 = Wigner3j(2 * self.Nl - 1, device=self.device)
        return self._w3j000

    @w3j000.setter
    def w3j000(self, value):
        self._w3j000 = value

    def __call__(self, Cl, bins, mode="deconvolution"):
        """Binned ells and mode-decoupled (or normalized) bandpowers
        (sht/mask_deconvolution.py:53-95)."""
        Cl = np.asarray(Cl)
        if len(Cl) < self.Nl:
            Clx = np.pad(Cl, (0, self.Nl - len(Cl)), "constant", constant_values=0)
        else:
            Clx = Cl[:self.Nl]
        Cb = bins @ Clx
        binned_ells = (bins @ np.arange(self.Nl)) / np.sum(bins, axis=1)
        if mode == "deconvolution":
            return binned_ells, self.decouple_Cls(self.get_Mbb_inv(bins), Cb)
        if mode == "normalization":
            return binned_ells, self.compute_normalization_factor() * Cb
        raise RuntimeError("Unknown mode. Must be 'deconvolution' or 'normalization'.")

    def compute_normalization_factor(self):
        """1 / [sum_l (2l+1) W_l / 4pi], the factor that leaves shot noise
        unchanged (sht/mask_deconvolution.py:97-104)."""
        ells = np.arange(len(self.W_l))
        return 1.0 / (np.sum(self.W_l * (2 * ells + 1)) / (4 * np.pi))

    def convolve_theory_Cls(self, Clt, bins, mode="deconvolution"):
        """(binned ells, window_matrix(bins, mode) . Clt[:Nl])
        (sht/mask_deconvolution.py:106-123)."""
        assert len(Clt) >= self.Nl, "Clt must be provided up to lmax"
        binned_ells = (bins @ np.arange(self.Nl)) / np.sum(bins, axis=1)
        return binned_ells, self.window_matrix(bins, mode) @ np.asarray(Clt)[:self.Nl]

    def window_matrix(self, bins, mode="deconvolution"):
        """Matrix from a theory C_l to binned, decoupled (or normalized)
        pseudo-spectra (sht/mask_deconvolution.py:125-153)."""
        if self.Mll is None:
            print("Computing the mode-coupling matrix...")
            self.Mll = self.get_M()
        if mode == "deconvolution":
            return self.get_Mbb_inv(bins) @ (bins @ self.Mll)
        if mode == "normalization":
            return self.compute_normalization_factor() * (bins @ self.Mll)
        raise RuntimeError("Unknown mode. Must be 'deconvolution' or 'normalization'.")

    def W(self, l, debug=False):
        """W_l, or with debug the full-sky window 4pi delta_{l0}
        (sht/mask_deconvolution.py:155-169)."""
        if debug:
            return 4 * np.pi if l == 0 else 0
        return self.W_l[l]

    def get_M(self, debug=False):
        """The (Nl, Nl) mode-coupling matrix (sht/mask_deconvolution.py:171-192),
        computed on the device; debug uses the full-sky window, for which M is
        the identity."""
        W = np.array([4 * np.pi]) if debug else self.W_l
        M = mode_coupling_device(self.Nl, W, self.device)
        return M.cpu().numpy()

    def get_Mbb_inv(self, bins):
        """Inverse of the bin-bin coupling bins . M . 1[bins > 0]^T
        (sht/mask_deconvolution.py:194-210)."""
        if self.Mll is None:
            print("Computing the mode-coupling matrix...")
            self.Mll = self.get_M()
        bins_no_wt = (np.asarray(bins) > 0).astype(np.float64)
        return np.linalg.inv((bins @ self.Mll) @ bins_no_wt.T)

    def binning_matrix(self, type="linear", start=0, step=16):
        """(Nbins, Nl) matrix averaging C_l over consecutive bins from `start`;
        'linear': width `step`, 'sqrt': width ceil(sqrt(4 l) + step) at the
        bin's first l; bins must end by lmax (sht/mask_deconvolution.py:212-238)."""
        Nl = self.lmax + 1
        if type == "linear":
            width = lambda ell: step            # noqa: E731
        elif type == "sqrt":
            width = lambda ell: int(np.ceil(np.sqrt(4.0 * ell) + step))  # noqa: E731
        else:
            raise RuntimeError("Unknown step type.")
        rows = []
        lo = start
        hi = lo + width(lo)
        while hi <= Nl:
            r = np.zeros(Nl)
            r[lo:hi] = 1.0 / float(hi - lo)
            rows.append(r)
            lo, hi = hi, hi + width(hi)
        return np.array(rows).reshape(len(rows), Nl)

    def decouple_Cls(self, Minv, Cb):
        """Minv . Cb (sht/mask_deconvolution.py:240-247)."""
        return Minv @ Cb
