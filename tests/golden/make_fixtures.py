"""Writes the committed fixtures under tests/golden/ (run in the build container only).

1. anastasio2D_vp.npz: the vp model of the reference's breast2D fixture
   (stride_examples/examples/breast2D/data/anastasio2D-TrueModel.h5, Mosaic h5
   protocol 0.1: dataset ``data`` 356 x 385 float32, spacing 0.5 mm, extra 50,
   absorbing 40). No h5py in this image: the dataset is stored contiguous and
   uncompressed, so its bytes are located by shape and value range and read with
   numpy (no code from the file is executed).
2. known_answers.json: the reference's logged scheme-selection lines
   (stride_examples/tutorials/03_acoustic_modelling.ipynb:305-310,
   stride_examples/examples/marmottant/00_acoustic_propagation.ipynb cell 19)
   with the inputs that produced them.
"""

import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF = '/root/reference'


def anastasio():
    path = os.path.join(REF, 'stride_examples/examples/breast2D/data/anastasio2D-TrueModel.h5')
    raw = open(path, 'rb').read()
    n = 356 * 385
    for off in range(0, len(raw) - 4 * n + 1, 4):
        a = np.frombuffer(raw, dtype='<f4', count=n, offset=off)
        if 1400 < a[0] < 1700 and np.all((a > 1400) & (a < 1700)):
            np.savez_compressed(os.path.join(HERE, 'anastasio2D_vp.npz'), vp=a.reshape(356, 385),
                                spacing=0.5e-3, extra=50, absorbing=40, offset=off)
            return
    raise RuntimeError('dataset not found')


def known_answers():
    kat = [
        dict(source='stride_examples/tutorials/03_acoustic_modelling.ipynb:305-310',
             shape=[200, 200], extra=[50, 50], absorbing=[40, 40], spacing=0.4e-3, dt=0.05e-6, nt=2000,
             vp=1500.0, f_centre=0.5e6, n_cycles=3, preferred_undersampling=None,
             lines=['Estimated bandwidth for the propagated wavelet 0.260-0.730 MHz',
                    'Spatial grid spacing (0.400 mm | 5.137 PPW) is below dispersion limit (0.411 mm | 5.000 PPW)',
                    'Time grid spacing (0.050 μs | 18%) is below OT2 limit (0.120 μs)',
                    'Selected undersampling level 6']),
        dict(source='stride_examples/examples/marmottant/00_acoustic_propagation.ipynb cell 19',
             shape=[100, 100, 100], extra=[30, 30, 30], absorbing=[20, 20, 20], spacing=0.05e-3, dt=0.005e-6,
             nt=1500, vp=1540.0, f_centre=2.0e6, n_cycles=6, preferred_undersampling=20,
             lines=['Estimated bandwidth for the propagated wavelet 1.600-2.400 MHz',
                    'Spatial grid spacing (0.050 mm | 12.833 PPW) is below dispersion limit (0.128 mm | 5.000 PPW)',
                    'Time grid spacing (0.005 μs | 15%) is below OT2 limit (0.012 μs)',
                    'Selected undersampling level 20',
                    'Selected time stepping scheme OT2']),
    ]
    with open(os.path.join(HERE, 'known_answers.json'), 'w') as f:
        json.dump(kat, f, indent=1, ensure_ascii=False)


if __name__ == '__main__':
    anastasio()
    known_answers()
