"""Configuration: the reference's decfg (dedalus/config.py:26-56) in Python 3.

Sections and defaults follow the reference; ``FFT.method`` defaults to
``'hip'``, the MI355X path of this package, which is the only method it
provides.
"""

import configparser
import os

decfg = configparser.ConfigParser(interpolation=None)

decfg.add_section('FFT')
decfg.set('FFT', 'method', 'hip')
decfg.set('FFT', 'dealiasing', '2/3 cython')

decfg.add_section('physics')
decfg.set('physics', 'use_tracer', 'False')
decfg.set('physics', 'boussinesq_direction', 'z')
decfg.add_section('forcing')

decfg.add_section('utils')
decfg.set('utils', 'loglevel', 'debug')
decfg.set('utils', 'loadplugins', 'False')
decfg.set('utils', 'pluginfilename', 'dedalus_plugins.py')

decfg.add_section('analysis')
decfg.set('analysis', 'snapshot_space', 'xspace')
decfg.set('analysis', 'snapshot_axis', 'z')
decfg.set('analysis', 'snapshot_index', 'middle')
decfg.set('analysis', 'snapshot_units', 'True')
decfg.set('analysis', 'snapshot_dpi', '100')
decfg.set('analysis', 'snapshot_cmap', 'Spectral_r')
decfg.set('analysis', 'powerspectrum_dpi', '100')

decfg.add_section('hip')
decfg.set('hip', 'device', '0')
decfg.set('hip', 'precision', 'float64')   # float32: FP32 transforms and products

decfg.read([os.path.expanduser('~/.dedalus/config'), 'dedalus.cfg'])
