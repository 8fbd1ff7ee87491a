"""Python layer of the TensorRay interface on MI355X: the autograd edge,
parameter manager, image conversions and losses of the reference's
pyTensorRay package, over the `TensorRay` extension module."""
