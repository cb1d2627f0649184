"""Abstract collision operator (reference ``src/collision/base.py:6-90``).

``__call__`` keeps the reference signature and its gain-only default
(:40, :52).  Both ``device="cpu"`` and ``device="gpu"`` compute on the HIP
device: a numpy input is copied to the device and the result copied back (a
fresh array, as the reference returns); a torch CUDA tensor stays resident and
a fresh tensor is returned.
"""
from abc import ABCMeta, abstractmethod


class BaseCollision(object, metaclass=ABCMeta):
    def __init__(self, config, spectral_mesh, **kwargs):
        self.config = config
        self.sm = spectral_mesh
        self.ndim = int(config["collision-model"]["dim"])
        assert self.ndim == self.sm.ndim, (
            "'spectral_mesh' has different velocity dimensions from collision model! "
            "This may be caused by using different config files. Please check the consistency.")
        self.options = kwargs
        self.load_parameters()
        self.perform_precomputation()
        self._input_shape = None

    def __call__(self, input_f, heat_bath=None, device="cpu"):
        if device not in ("cpu", "gpu", "cuda", "hip"):
            raise ValueError("unknown device %r" % (device,))
        self._input_shape = tuple(input_f.shape)
        return self.collide(input_f, with_loss=self.call_with_loss, heat_bath=heat_bath)

    call_with_loss = False

    @abstractmethod
    def perform_precomputation(self):
        pass

    @abstractmethod
    def load_parameters(self):
        pass

    @abstractmethod
    def collide(self, input_f, with_loss=False, heat_bath=None):
        pass

    @abstractmethod
    def laplacian(self, input_f):
        pass
