n1=101 o1=0 d1=10
n2=201 o2=0 d2=10
n3=3 o3=0 d3=10
esize=4
data_format="xdr_float"
in=./smooth_layers.H@
