"""Per-element conductivities (torchcor/core/conductivity.py:4-69)."""
import torch


class Conductivity:
    """Region-wise longitudinal/transverse conductivities.

    ``add(region_ids, il, it)`` sets the monodomain values directly and
    drops the intra/extracellular fields; ``add(..., el, et)`` stores the
    harmonic means ``il*el*(1/(il+el))``.  Regions never added keep 0.
    """

    def __init__(self, regions, dtype):
        self.regions = regions
        self.dtype = dtype
        z = lambda: torch.zeros(regions.shape[0], dtype=dtype, device=regions.device)  # noqa: E731
        self.sigma_il, self.sigma_it, self.sigma_el, self.sigma_et = z(), z(), z(), z()
        self.sigma_l, self.sigma_t = z(), z()

    def add(self, region_ids, il, it, el=None, et=None):
        if el is None and et is None:
            self.sigma_il = self.sigma_it = self.sigma_el = self.sigma_et = None
            for rid in region_ids:
                m = self.regions == rid
                self.sigma_l[m] = il
                self.sigma_t[m] = it
        else:
            l = il * el * (1 / (il + el))
            t = it * et * (1 / (it + et))
            for rid in region_ids:
                m = self.regions == rid
                self.sigma_il[m] = il
                self.sigma_it[m] = it
                self.sigma_el[m] = el
                self.sigma_et[m] = et
                self.sigma_l[m] = l
                self.sigma_t[m] = t

    @staticmethod
    def _tensor(l, t, fibres):
        eye = torch.eye(3, dtype=fibres.dtype, device=fibres.device)
        return t.view(-1, 1, 1) * eye + (l - t).view(-1, 1, 1) * (fibres.unsqueeze(2) @ fibres.unsqueeze(1))

    def calculate_sigma(self, fibres):
        """(sigma_i, sigma_e, sigma_m), each (E,3,3) or None (conductivity.py:49-69)."""
        fibres = fibres.to(self.dtype)
        if self.sigma_il is None:
            si = se = None
        else:
            si = self._tensor(self.sigma_il, self.sigma_it, fibres)
            se = self._tensor(self.sigma_el, self.sigma_et, fibres)
        return si, se, self._tensor(self.sigma_l, self.sigma_t, fibres)
