"""Loader of libvpichip.so (placeholder until the HIP library lands)."""
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libvpichip.so")
