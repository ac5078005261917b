from .matrix import *
