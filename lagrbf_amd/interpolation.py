"""LagRBFInterpolator, mirroring python_api/lagrbfpy/interpolation.py:35-116."""
from .context import LagRBFCtx


class LagRBFInterpolator(LagRBFCtx):
    """PU-RBF interpolation from data points to arbitrary points on the GPU.

    As in the reference, the RBF systems carry no polynomial augmentation.
    """

    def interpolate(self, data, points, output, fieldname="fields"):
        """fill_data -> fill_points -> interpolate -> output -> reset_coefficients
        (interpolation.py:97-115)."""
        self.fill_data(coords=data.coords, values=data.values, pidx=data.pidx)
        self.fill_points(coords=points.coords, values=points.values, pidx=points.pidx)
        self.interpolate_c()
        self.output(coords=output.coords, fvals=output.values, pidx=output.pidx, fieldname=fieldname, is_data=False)
        self.reset_coefficients()
